// Shared device/host helpers of libmipt_hip: status handling, bit tricks,
// the Philox4x32-10 counter-based generator and the device Haar U(4) sampler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace mipt_hip
{

// ------------------------------------------------------------ status codes
// 0 = success; -1 invalid argument; -2 invalid branch norms / zero state;
// -3 invalid conditional probability; -4 invalid branch norm (the codes of
// cpp/src/cuda/free_energy_measure.cu:137-238); -5 unsupported request;
// -6 out of memory; positive = hipError_t; 100000 + status = rocSOLVER /
// rocBLAS failure; 102000 + info = FP64 eigensolver did not converge.
enum Status : int
{
    OK = 0,
    E_ARG = -1,
    E_NORM = -2,
    E_PROB = -3,
    E_BRANCH = -4,
    E_UNSUPPORTED = -5,
    E_NOMEM = -6,
    E_SOLVER = 100000,
    E_EIG = 102000,
};

void set_error(const char *fmt, ...);
const char *last_error();

#define MIPT_HIP_CHECK(expr)                                                          \
    do                                                                                \
    {                                                                                 \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
        {                                                                             \
            ::mipt_hip::set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__, \
                                  hipGetErrorString(_e));                             \
            return static_cast<int>(_e);                                              \
        }                                                                             \
    } while (0)

#define MIPT_HIP_TRY(expr)                 \
    do                                     \
    {                                      \
        int _s = (expr);                   \
        if (_s != 0) return _s;            \
    } while (0)

// ------------------------------------------------------------ bit helpers
// Scatter the low bits of `v` into the set bits of `mask` (software pdep).
__host__ __device__ inline std::uint64_t pdep64(std::uint64_t v, std::uint64_t mask)
{
    std::uint64_t out = 0;
    for (std::uint64_t bit = 1; mask; bit <<= 1)
    {
        const std::uint64_t low = mask & (~mask + 1);
        if (v & bit) out |= low;
        mask ^= low;
    }
    return out;
}

// Gather the bits of `v` selected by `mask` into the low bits (software pext).
__host__ __device__ inline std::uint64_t pext64(std::uint64_t v, std::uint64_t mask)
{
    std::uint64_t out = 0;
    for (std::uint64_t bit = 1; mask; bit <<= 1)
    {
        const std::uint64_t low = mask & (~mask + 1);
        if (v & low) out |= bit;
        mask ^= low;
    }
    return out;
}

// Next submask of `mask` after `x` in increasing order (x must be a submask).
__host__ __device__ inline std::uint32_t next_submask(std::uint32_t x, std::uint32_t mask)
{
    return ((x | ~mask) + 1u) & mask;
}

// ------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11).  Counter words: {slot, item, layer, stream}; key = trajectory seed.
struct U32x4
{
    std::uint32_t x, y, z, w;
};

__host__ __device__ inline U32x4 philox4x32_10(std::uint32_t k0, std::uint32_t k1, U32x4 c)
{
#pragma unroll
    for (int round = 0; round < 10; ++round)
    {
        const std::uint64_t p0 = static_cast<std::uint64_t>(0xD2511F53u) * c.x;
        const std::uint64_t p1 = static_cast<std::uint64_t>(0xCD9E8D57u) * c.z;
        const std::uint32_t hi0 = static_cast<std::uint32_t>(p0 >> 32), lo0 = static_cast<std::uint32_t>(p0);
        const std::uint32_t hi1 = static_cast<std::uint32_t>(p1 >> 32), lo1 = static_cast<std::uint32_t>(p1);
        c = U32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

constexpr std::uint32_t STREAM_GATE = 0x47415445u;
constexpr std::uint32_t STREAM_FLAG = 0x464c4147u;
constexpr std::uint32_t STREAM_MEAS = 0x4d454153u;

__host__ __device__ inline double u53(std::uint32_t a, std::uint32_t b)
{
    return (static_cast<double>(a >> 5) * 67108864.0 + static_cast<double>(b >> 6)) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline void philox_uniform_pair(std::uint64_t seed, std::uint32_t layer, std::uint32_t item,
                                                    std::uint32_t slot, std::uint32_t stream, double &u, double &v)
{
    const U32x4 r = philox4x32_10(static_cast<std::uint32_t>(seed & 0xffffffffu),
                                  static_cast<std::uint32_t>(seed >> 32), U32x4{slot, item, layer, stream});
    u = u53(r.x, r.y);
    v = u53(r.z, r.w);
}

// Haar U(4) by modified Gram-Schmidt on complex normal columns (the sampler of
// cpp/include/mipt/circuits/haar.hpp:39-85), normals by Box-Muller on the
// counter-based stream keyed by (seed, layer, bond).  Row-major FP64 output,
// basis index 2*b(q1) + b(q0).  Re before Im, cos before sin.
__host__ __device__ inline void haar_u4_philox(std::uint64_t seed, std::uint32_t layer, std::uint32_t bond,
                                               double qr[16], double qi[16])
{
    std::uint32_t slot = 0;
    for (int col = 0; col < 4; ++col)
    {
        double vr[4], vi[4];
        for (int row = 0; row < 4; ++row)
        {
            double u, v;
            philox_uniform_pair(seed, layer, bond, slot++, STREAM_GATE, u, v);
            const double r = sqrt(-2.0 * log(1.0 - u));
            const double a = 6.283185307179586476925286766559 * v;
            vr[row] = r * cos(a);
            vi[row] = r * sin(a);
        }
        for (int prev = 0; prev < col; ++prev)
        {
            double dr = 0.0, di = 0.0;  // dot = sum conj(q) * v
            for (int row = 0; row < 4; ++row)
            {
                const double ar = qr[row * 4 + prev], ai = qi[row * 4 + prev];
                dr += ar * vr[row] + ai * vi[row];
                di += ar * vi[row] - ai * vr[row];
            }
            for (int row = 0; row < 4; ++row)
            {
                const double ar = qr[row * 4 + prev], ai = qi[row * 4 + prev];
                vr[row] -= ar * dr - ai * di;
                vi[row] -= ar * di + ai * dr;
            }
        }
        double n2 = 0.0;
        for (int row = 0; row < 4; ++row) n2 += vr[row] * vr[row] + vi[row] * vi[row];
        const double nrm = sqrt(n2);
        if (nrm < 1e-14)
        {
            --col;
            continue;
        }
        for (int row = 0; row < 4; ++row)
        {
            qr[row * 4 + col] = vr[row] / nrm;
            qi[row * 4 + col] = vi[row] / nrm;
        }
    }
}

// Brickwork layout (cpp/include/mipt/circuits/haar.hpp:226-243).
__host__ __device__ inline int layer_bond_count(int n, int layer, int closed)
{
    const int start = layer & 1;
    const int stop = (closed && start == 1 && n > 2) ? n : (n - 1);
    return stop > start ? (stop - start + 1) / 2 : 0;
}

__host__ __device__ inline void layer_bond(int n, int layer, int b, int &q0, int &q1)
{
    q0 = (layer & 1) + 2 * b;
    q1 = (q0 + 1) % n;
}

// Bond layouts by geometry.  gx = 0: the brickwork chain above (two layers
// per period).  gx > 0: an n = gx * gy grid, site r*gx + c, four bond colours
// per time step (layer mod 4), the layering of the reference's 2D circuit
// (cpp/include/mipt/circuits/mms.hpp:159-330) with Haar bonds:
//   0: horizontal bonds (c, c+1) from even columns, row by row;
//   1: from odd columns, plus the row closure (r*gx + gx-1, r*gx) when closed
//      and gx is even (an odd gx would give the last site two bonds);
//   2, 3: vertical bonds (r, r+1) in column c from rows
//      ((c mod 2) + colour - 2) mod 2 in steps of 2, column by column.
__host__ __device__ inline int period_layers(int gx) { return gx > 0 ? 4 : 2; }

__host__ __device__ inline int vertical_start(int c, int colour) { return ((c & 1) + colour - 2) & 1; }

__host__ __device__ inline int geo_bond_count(int n, int gx, int layer, int closed)
{
    if (gx <= 0) return layer_bond_count(n, layer, closed);
    const int gy = n / gx, colour = layer & 3;
    if (colour < 2)
    {
        const int per_row = (gx - colour) / 2 + ((colour == 1 && closed && (gx % 2) == 0 && gx > 2) ? 1 : 0);
        return per_row * gy;
    }
    int total = 0;
    for (int c = 0; c < gx; ++c)
    {
        const int s = vertical_start(c, colour);
        if (gy - 2 >= s) total += (gy - 2 - s) / 2 + 1;
    }
    return total;
}

__host__ __device__ inline void geo_bond(int n, int gx, int layer, int b, int closed, int &q0, int &q1)
{
    if (gx <= 0)
    {
        layer_bond(n, layer, b, q0, q1);
        return;
    }
    const int gy = n / gx, colour = layer & 3;
    if (colour < 2)
    {
        const int regular = (gx - colour) / 2;
        const int per_row = regular + ((colour == 1 && closed && (gx % 2) == 0 && gx > 2) ? 1 : 0);
        const int r = b / per_row, k = b % per_row;
        if (k < regular)
        {
            q0 = r * gx + colour + 2 * k;
            q1 = q0 + 1;
        }
        else
        {
            q0 = r * gx + gx - 1;
            q1 = r * gx;
        }
        return;
    }
    for (int c = 0; c < gx; ++c)
    {
        const int s = vertical_start(c, colour);
        const int cnt = (gy - 2 >= s) ? (gy - 2 - s) / 2 + 1 : 0;
        if (b < cnt)
        {
            const int r = s + 2 * b;
            q0 = r * gx + c;
            q1 = (r + 1) * gx + c;
            return;
        }
        b -= cnt;
    }
    q0 = q1 = -1;  // unreachable for b < geo_bond_count
}

// Gates of layers [0, t) of a trajectory (the gate stream's layer offset).
__host__ __device__ inline std::uint64_t geo_gate_prefix(int n, int gx, int closed, int t)
{
    const int P = period_layers(gx);
    std::uint64_t per = 0, part = 0;
    for (int j = 0; j < P; ++j)
    {
        const int c = geo_bond_count(n, gx, j, closed);
        per += static_cast<std::uint64_t>(c);
        if (j < t % P) part += static_cast<std::uint64_t>(c);
    }
    return static_cast<std::uint64_t>(t / P) * per + part;
}

} // namespace mipt_hip
