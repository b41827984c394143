// Batched SST-2003 local state and wall-law evaluations on the GPU (host
// arrays in and out).  Layouts follow the C ABI in include/nodals_hip.h.
#pragma once

namespace nodals {

void sstLocalBatch(long long n, const double* dist, const double* xy, double radius, double nu, const double* gu,
                   const double* k, const double* omega, const double* gk, const double* gw, double kFloor,
                   double omegaFloor, int deviatoric, double* out, int* flags);
void spaldingBatch(long long n, const double* slip, const double* y, double nu, double* out, int* ok);
void profile5Batch(long long n, const double* U, const double* Us, const double* y, double nu, double* out,
                   int* valid);
// The same fit with no allocation: staging in a fixed device array of the
// library (at most 65536 faces per call; calls that overlap on one GPU race).
void profile5BatchNoAlloc(long long n, const double* U, const double* Us, const double* y, double nu, double* out,
                          int* valid);
void fitTangentialBatch(long long n, const double* samples, const double* y, const double* nrm, const double* pic,
                        double nu, int count, int zeroLimit, double* out, int* valid);
void omegaWallBatch(long long n, const double* in, const int* shared, double* out);

}  // namespace nodals
