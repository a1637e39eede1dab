/*
 * interface.h -- C++ interface of librbergomi_mi355x.so (C++ linkage, std::vector).
 *
 * Same declarations as the reference's rBergLib/interface.h:23-65 so that code compiled
 * against it links unchanged: ResultSimple and the four Compute* wrappers.  The reference's
 * test stubs ParSimple/stupid (:39-49, :67-68) and Result2ResultSimple (which takes the
 * reference-internal Result/ParamTot types) have no counterpart.
 *
 * singleThread and multiThread variants run the same GPU pipeline; numThreads is echoed in
 * ResultSimple::numThreads only.  Errors throw std::runtime_error (the reference terminates).
 */
#pragma once

#include <stdint.h>

#include <vector>

typedef std::vector<double> Vector;

/* Output index i = h * (number of (T, K) pairs) + t for parameter triple h and option t; the
 * per-index echo vectors say which inputs produced entry i.  The member order is the ABI. */
struct ResultSimple {
    Vector price;  /* Monte Carlo price of entry i */
    Vector iv;     /* Black implied volatility of price[i] */
    Vector H;      /* H of entry i */
    Vector rho;    /* rho of entry i (rho before eta, as the reference orders them) */
    Vector eta;    /* eta of entry i */
    Vector T;      /* maturity of entry i */
    Vector K;      /* strike of entry i */
    double xi;
    Vector stat;   /* Monte Carlo standard error of price[i] */
    int N;
    long M;
    int numThreads;
};

/* rBergLib/interface.h:51-52 -- plain Monte Carlo */
ResultSimple ComputeIV_singleThread(double xi, Vector H, Vector eta, Vector rho, Vector T, Vector K,
                                    int N, long M, std::vector<uint64_t> seed);
/* rBergLib/interface.h:54-55 */
ResultSimple ComputeIV_multiThread(double xi, Vector H, Vector eta, Vector rho, Vector T, Vector K,
                                   int N, long M, int numThreads, std::vector<uint64_t> seed);
/* rBergLib/interface.h:58-59 -- Romano-Touzi */
ResultSimple ComputeIVRT_singleThread(double xi, Vector H, Vector eta, Vector rho, Vector T,
                                      Vector K, int N, long M, std::vector<uint64_t> seed);
/* rBergLib/interface.h:61-62 */
ResultSimple ComputeIVRT_multiThread(double xi, Vector H, Vector eta, Vector rho, Vector T,
                                     Vector K, int N, long M, int numThreads,
                                     std::vector<uint64_t> seed);
