// Device builds of numerics whose device code differs from the host build (the fast square root)
// or whose exactness argument is worth checking on the GPU (the limiter's four-face minimum), for
// the GPU tests to compare with the library routines bit for bit.  Test infrastructure only; built
// in-tree by __graft_entry__.build() (tests/native/libdevice_numerics.so).
#include <hip/hip_runtime.h>

#include "../../pyhype_amd/csrc/phx_numerics.hpp"

namespace {

__global__ void k_sqrt(const double* x, double* fast, double* ref, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        fast[i] = phx::sqrt_rn(x[i]);
        ref[i] = sqrt(x[i]);
    }
}

__global__ void k_min4(const double* p, double* fast, double* chain, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double* q = p + 4 * (long)i;
        fast[i] = phx::limiter_min4<false>(q[0], q[1], q[2], q[3]);
        chain[i] = phx::limiter_min4<true>(q[0], q[1], q[2], q[3]);
    }
}

}  // namespace

extern "C" {
// device arrays of n doubles (p: 4n); 0 on success, else the HIP error code
int dn_sqrt(const double* x, double* fast, double* ref, int n, void* stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_sqrt, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, fast, ref, n);
    return (int)hipGetLastError();
}
int dn_min4(const double* p, double* fast, double* chain, int n, void* stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_min4, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, fast, chain, n);
    return (int)hipGetLastError();
}
}
