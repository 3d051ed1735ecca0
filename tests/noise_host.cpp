// Host build of the product's noise evaluation (nixis_amd/csrc/nx_noise_dev.h)
// for CPU tests: the same region logic, hashing and FP32 arithmetic as the
// kernels, compiled by the host compiler (-ffp-contract=off), with the lane
// masks of the region logic reduced to all-or-nothing masks of one lane.
// The tables are laid out as the kernels' loader lays them out in LDS.
#include <cstdint>
#include <cstring>

#include "nx_noise_dev.h"

namespace {

// Kernel-argument tables of opensimplex.py:51-60 and init() output, then the
// LDS layout (padded bytes, gradient byte offsets, 64-bit gradient words).
int build_tables(const int32_t *perm, const int32_t *pgi, nx::NoiseTables &T)
{
    uint8_t pb[256], gb[256];
    for (int i = 0; i < 256; ++i) {
        if (perm[i] < 0 || perm[i] > 255 || pgi[i] < 0 || pgi[i] > 69 || pgi[i] % 3) return -1;
        pb[i] = (uint8_t)perm[i];
        gb[i] = (uint8_t)(8 * (pgi[i] / 3));
    }
    for (int i = 0; i < nx::kWrap; ++i) {
        T.perm[i] = pb[(i + 255) & 255];
        T.grad[i] = gb[(i + 255) & 255];
    }
    for (int g = 0; g < 24; ++g) {
        const int s = g / 3, k = g % 3;
        const int gx = (k == 0 ? 11 : 4) * ((s & 1) ? 1 : -1);
        const int gy = (k == 1 ? 11 : 4) * ((s & 2) ? -1 : 1);
        const int gz = (k == 2 ? 11 : 4) * ((s & 4) ? -1 : 1);
        const uint32_t lo = nx::half_of_small(gx) | (nx::half_of_small(gy) << 16);
        const float fz = (float)gz;
        uint32_t hi;
        std::memcpy(&hi, &fz, 4);
        T.g16[g] = (uint64_t)lo | ((uint64_t)hi << 32);
    }
    static const nx::ExtLut lut = nx::make_ext_lut();
    for (int i = 0; i < nx::kExtIndex; ++i) T.ext[i] = lut.e[i];
    return 0;
}

}  // namespace

// noise3d at n points (xyz [n][3]): FP32 values and the selection records
// (xsb, ysb, zsb, region | extra-vertex offsets, gradient checksum); returns
// the number of points outside the exact domain, or -1 for bad tables.
extern "C" int64_t host_noise3(const double *xyz, int64_t n, const int32_t *perm,
                               const int32_t *pgi, float *val, int32_t *sel)
{
    static nx::NoiseTables T;
    if (build_tables(perm, pgi, T)) return -1;
    int64_t out = 0;
    for (int64_t i = 0; i < n; ++i) {
        bool bad = false;
        nx::NoiseSel s;
        val[i] = nx::noise3<true>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], T, bad, &s);
        std::memcpy(sel + 5 * i, s.v, sizeof(s.v));
        out += bad;
    }
    return out;
}
