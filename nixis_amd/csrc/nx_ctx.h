// nx_ctx.h -- the context object behind nx_ctx* and HIP error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "nx_internal.h"

// Bits of the per-context sticky error word written by kernels.
#define NX_ERR_RANGE_BIT 1
#define NX_ERR_DIV0_BIT 2

constexpr int kKeySlots = 256;

struct nx_ctx {
    int device = 0;
    int num_cus = 256;
    int *d_err = nullptr;                  // device word, bits NX_ERR_*
    unsigned long long *d_keys = nullptr;  // kKeySlots x 2 reduction keys
    std::atomic<unsigned> key_slot{0};
    std::mutex mu;
};

#define NX_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::nx::set_error(NX_EHIP, "%s failed: %s (%s:%d)", #call,               \
                                   hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

namespace nx {

int sample_octaves(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                   const int32_t *pgi, int n_oct, const double *rough, const double *strength,
                   double radius, double *elev, int accumulate, float *dbg_noise,
                   int32_t *dbg_sel, hipStream_t stream);
int noise3d(nx_ctx *ctx, const double *x, const double *y, const double *z, int64_t n,
            const int32_t *perm, const int32_t *pgi, double *out, hipStream_t stream);
int rescale(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper, double *out,
            double *minmax, hipStream_t stream);
size_t erosion_state_bytes(int64_t V);
int erosion_load(nx_ctx *ctx, void *state, int64_t V, const double *h, const double *w,
                 const double *s, hipStream_t stream);
int erosion_run(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const float *elen,
                int iters, double rain, hipStream_t stream);
int erosion_store(nx_ctx *ctx, const void *state, int64_t V, double *h, double *w, double *s,
                  hipStream_t stream);
int erosion_mass(nx_ctx *ctx, const void *state, int64_t V, double *out, hipStream_t stream);

}  // namespace nx
