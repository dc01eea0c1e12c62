// Batched FP64 Hermitian eigenvalues and von Neumann entropies of reduced
// density matrices: many matrices of one size in one launch.
//
// Reference behaviour (paths relative to the reference repository root):
//   cpp/src/cuda/tmi_cuda_svd.cu:663-798    device entropy: heevd / gesvd ladder (cuSOLVER)
//   cpp/include/mipt/tmi/linalg.hpp:524-583 host Hermitian entropy (heevd or Jacobi)
//   cpp/include/mipt/tmi/linalg.hpp:125-141 entropy_from_eigenvalue / finish_entropy cut-offs
//
// gfx950 design: one workgroup per matrix, no host round trips between steps.
//
// Stage 1 (hetrd_kernel) reduces a Hermitian m x m matrix to a real symmetric
// tridiagonal (d, e) by Householder reflections in panels of 32 columns: the
// lower-panel form of LAPACK's zhetrd/zlatrd.  Per column: the column is
// brought up to date with the panel's earlier reflectors, its reflector is
// formed (one norm reduction), and w = A v comes from one Hermitian
// matrix-vector product over the stored lower triangle of the trailing
// matrix, corrected by the panel terms.  After the panel the trailing
// triangle takes the panel's rank-2 terms A -= V W^H + W V^H at once.  The
// matrix-vector products dominate (8 m^3 / 3 bytes read per matrix, the
// HBM-bound part).  They are tiled 16 x 32: lane (row, column group) reads 8
// consecutive-column entries of one row, so a wave instruction reads four
// 256-byte column segments; the lane forms the row sum of its row and the
// conjugate column sums of its 8 columns.  Row sums go to the wave's own LDS
// vector, column sums are reduced over the 16 row lanes once per column
// strip, and the per-wave vectors are added in a fixed order at the end, so
// results do not depend on scheduling.  Column strips are dealt to waves in
// a snake order (strip J and the mirror strip share a wave) for balance.
//
// Stage 2 (tridiag_entropy_kernel) finds every eigenvalue by bisection on
// Sturm counts of (d, e) with LAPACK's pivot guard, then sums
// -lambda log2 lambda over lambda > 1e-15 in a fixed order and zeroes
// |S| < 1e-12, the host cut-offs of linalg.hpp:125-141.
//
// All arithmetic is FP64.  The matrices are consumed (overwritten by the
// reflectors).
#include "batch.hpp"

#include <cfloat>

namespace mipt_hip
{
namespace
{

constexpr int EIG_NB = 32;    // panel width
constexpr int EIG_TILE = 32;  // matrix-vector tile edge

struct HetrdArgs
{
    double2 *a;            // matrices (column-major view, lower triangle read)
    std::size_t a_stride;  // elements between matrices
    double2 *w;            // panel workspace, m x EIG_NB per matrix
    std::size_t w_stride;
    double *d, *e;         // m per matrix (e: m - 1 used)
    int n;
};

// Sum over the workgroup, fixed order: wave butterflies, then the waves in
// order.  `red` holds NT/64 doubles.
template <int NT>
__device__ double block_sum(double v, double *red)
{
    constexpr int NW = NT / 64;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();  // earlier readers of red are done
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[w];
    return s;
}

template <int NT>
__global__ void __launch_bounds__(NT) hetrd_kernel(HetrdArgs g)
{
    constexpr int NW = NT / 64;
    extern __shared__ double2 lds[];
    const int n = g.n;
    const std::size_t nn = static_cast<std::size_t>(n);
    double2 *__restrict__ A = g.a + blockIdx.x * g.a_stride;
    double2 *__restrict__ W = g.w + blockIdx.x * g.w_stride;
    double *__restrict__ dout = g.d + blockIdx.x * nn;
    double *__restrict__ eout = g.e + blockIdx.x * nn;
    const int ysz = max(NW * n, 2 * EIG_TILE * EIG_NB);
    double2 *vs = lds;              // [n] current reflector, index r - s
    double2 *yp = lds + n;          // [NW][n] per-wave products; panel staging of the rank-2 update
    double2 *vrow = yp + ysz;       // [NB] row i of the panel's reflectors
    double2 *wrow = vrow + EIG_NB;  // [NB] row i of the panel's W
    double2 *tv = wrow + EIG_NB;    // [2 NB] W^H v, V^H v
    double *red = reinterpret_cast<double *>(tv + 2 * EIG_NB);  // [NW]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rr = lane & 31, h = lane >> 5;    // rank-2 update tiles: 32 rows x 2 x 16 columns
    const int q16 = lane & 15, g8 = lane >> 4;  // matrix-vector tiles: 16 rows x 4 x 8 columns
    const double2 zero = make_double2(0.0, 0.0);

    // columns below 1e-120 of the largest diagonal entry are taken as zero
    // (identity reflector): past the numerical rank of an RDM the trailing
    // columns are rounding noise that shrinks at every step, and its
    // reflector would otherwise underflow (backward error < 1e-120 scale)
    double dmax = 0.0;
    for (int r = tid; r < n; r += NT) dmax = fmax(dmax, fabs(A[r + r * nn].x));
    for (int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
    if (lane == 0) red[wv] = dmax;
    __syncthreads();
    for (int w = 0; w < NW; ++w) dmax = fmax(dmax, red[w]);
    const double tiny2 = (1.0e-120 * dmax) * (1.0e-120 * dmax);

    for (int p0 = 0; p0 < n; p0 += EIG_NB)
    {
        const int nbp = min(EIG_NB, n - p0);
        for (int j = 0; j < nbp; ++j)
        {
            const int i = p0 + j;
            // column i takes the panel's earlier reflectors:
            // A(i:n, i) -= V(i:n, 0:j) conj(W(i, 0:j))^T + W(i:n, 0:j) conj(V(i, 0:j))^T
            if (tid < j)
            {
                vrow[tid] = A[i + (p0 + tid) * nn];
                wrow[tid] = W[i + tid * nn];
            }
            __syncthreads();
            for (int r = i + tid; r < n; r += NT)
            {
                double ax = A[r + i * nn].x, ay = A[r + i * nn].y;
                for (int k = 0; k < j; ++k)
                {
                    const double2 vk = A[r + (p0 + k) * nn], wk = W[r + k * nn];
                    const double2 wi = wrow[k], vi = vrow[k];
                    ax -= vk.x * wi.x + vk.y * wi.y + wk.x * vi.x + wk.y * vi.y;
                    ay -= vk.y * wi.x - vk.x * wi.y + wk.y * vi.x - wk.x * vi.y;
                }
                if (r == i) ay = 0.0;
                A[r + i * nn] = make_double2(ax, ay);
            }
            __syncthreads();
            if (tid == 0) dout[i] = A[i + i * nn].x;
            if (i == n - 1) break;

            // reflector of A(s:n, i), s = i + 1: beta real, v = [1; x / (alpha - beta)]
            const int s = i + 1, len = n - s;
            const double2 alpha = A[s + i * nn];
            double part = 0.0;
            for (int r = s + 1 + tid; r < n; r += NT)
            {
                const double2 x = A[r + i * nn];
                part += x.x * x.x + x.y * x.y;
            }
            const double xn2 = block_sum<NT>(part, red);
            double beta, taur, taui, sx, sy;
            const double s2 = alpha.x * alpha.x + alpha.y * alpha.y + xn2;
            if ((xn2 == 0.0 && alpha.y == 0.0) || s2 <= tiny2)
            {
                // identity reflector; |alpha| is the off-diagonal (a diagonal
                // phase makes it real without changing the spectrum)
                beta = sqrt(alpha.x * alpha.x + alpha.y * alpha.y);
                taur = taui = 0.0;
                sx = sy = 0.0;
            }
            else
            {
                beta = -copysign(sqrt(s2), alpha.x);
                taur = (beta - alpha.x) / beta;
                taui = -alpha.y / beta;
                // 1 / (alpha - beta) with the components scaled to [-1, 1]
                const double dx = alpha.x - beta, dy = alpha.y, t = fmax(fabs(dx), fabs(dy));
                const double ux = dx / t, uy = dy / t, uu = (ux * ux + uy * uy) * t;
                sx = ux / uu;
                sy = -uy / uu;
            }
            for (int r = s + tid; r < n; r += NT)
            {
                double2 x = make_double2(1.0, 0.0);
                if (r > s)
                {
                    const double2 a = A[r + i * nn];
                    x = make_double2(a.x * sx - a.y * sy, a.x * sy + a.y * sx);
                }
                A[r + i * nn] = x;
                vs[r - s] = x;
            }
            if (tid == 0) eout[i] = beta;
            for (int t = lane; t < len; t += 64) yp[wv * n + t] = zero;
            __syncthreads();

            // y = A(s:n, s:n) v over the stored lower triangle: strips of 32
            // columns, row blocks of 16; lane (row q, column group g) holds 8
            // columns of one row
            const int ntl = (len + EIG_TILE - 1) / EIG_TILE;
            const int nrb = (len + 15) / 16;
            for (int J = 0; J < ntl; ++J)
            {
                const int m2 = J % (2 * NW);
                if ((m2 < NW ? m2 : 2 * NW - 1 - m2) != wv) continue;
                double cx[8], cy[8];
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) cx[cc] = cy[cc] = 0.0;
                const int lc0 = EIG_TILE * J + 8 * g8;
                const double2 *colp = A + s + (s + min(lc0, len - 1)) * nn;
                // row block I's entries; the next block's loads are issued
                // before this block's arithmetic (two blocks in flight)
                double2 a[8], an[8];
                {
                    const int lr = 16 * (2 * J) + q16;
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc)
                    {
                        const int lc = lc0 + cc;
                        a[cc] = (lr < len && lc < len && lr >= lc) ? colp[min(lr, len - 1) + cc * nn] : zero;
                    }
                }
                for (int I = 2 * J; I < nrb; ++I)
                {
                    const int lr = 16 * I + q16;
                    const bool rv = lr < len;
                    const double2 vi = rv ? vs[lr] : zero;
                    const int lrn = lr + 16;
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc)
                    {
                        const int lc = lc0 + cc;
                        an[cc] = (lrn < len && lc < len && lrn >= lc) ? colp[min(lrn, len - 1) + cc * nn] : zero;
                    }
                    double rx = 0.0, ry = 0.0;
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc)
                    {
                        const int lc = lc0 + cc;
                        const double2 vc = vs[min(lc, len - 1)];
                        if (lr == lc)
                        {
                            rx += a[cc].x * vc.x;
                            ry += a[cc].x * vc.y;
                        }
                        else
                        {
                            rx += a[cc].x * vc.x - a[cc].y * vc.y;
                            ry += a[cc].x * vc.y + a[cc].y * vc.x;
                            cx[cc] += a[cc].x * vi.x + a[cc].y * vi.y;
                            cy[cc] += a[cc].x * vi.y - a[cc].y * vi.x;
                        }
                    }
                    rx += __shfl_xor(rx, 16, 64);
                    ry += __shfl_xor(ry, 16, 64);
                    rx += __shfl_xor(rx, 32, 64);
                    ry += __shfl_xor(ry, 32, 64);
                    if (g8 == 0 && rv)
                    {
                        double2 t = yp[wv * n + lr];
                        t.x += rx;
                        t.y += ry;
                        yp[wv * n + lr] = t;
                    }
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) a[cc] = an[cc];
                }
#pragma unroll
                for (int cc = 0; cc < 8; ++cc)
                    for (int off = 8; off > 0; off >>= 1)
                    {
                        cx[cc] += __shfl_xor(cx[cc], off, 64);
                        cy[cc] += __shfl_xor(cy[cc], off, 64);
                    }
                if (q16 == 0)
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc)
                    {
                        const int lc = lc0 + cc;
                        if (lc < len)
                        {
                            double2 t = yp[wv * n + lc];
                            t.x += cx[cc];
                            t.y += cy[cc];
                            yp[wv * n + lc] = t;
                        }
                    }
            }
            __syncthreads();

            // panel corrections: tv[k] = W(s:n, k)^H v, tv[j + k] = V(s:n, k)^H v
            for (int q = wv; q < 2 * j; q += NW)
            {
                const double2 *col = q < j ? W + q * nn : A + (p0 + q - j) * nn;
                double pr = 0.0, pi = 0.0;
                for (int r = s + lane; r < n; r += 64)
                {
                    const double2 x = col[r], v = vs[r - s];
                    pr += x.x * v.x + x.y * v.y;
                    pi += x.x * v.y - x.y * v.x;
                }
                for (int off = 32; off > 0; off >>= 1)
                {
                    pr += __shfl_xor(pr, off, 64);
                    pi += __shfl_xor(pi, off, 64);
                }
                if (lane == 0) tv[q] = make_double2(pr, pi);
            }
            __syncthreads();

            // w = tau (y - V tv[0:j] - W tv[j:2j]);  w += (-tau/2 w^H v) v
            double dr = 0.0, di = 0.0;
            for (int r = s + tid; r < n; r += NT)
            {
                const int lr = r - s;
                double yx = 0.0, yy = 0.0;
                for (int w = 0; w < NW; ++w)
                {
                    yx += yp[w * n + lr].x;
                    yy += yp[w * n + lr].y;
                }
                for (int k = 0; k < j; ++k)
                {
                    const double2 vk = A[r + (p0 + k) * nn], wk = W[r + k * nn], t1 = tv[k], t2 = tv[j + k];
                    yx -= vk.x * t1.x - vk.y * t1.y + wk.x * t2.x - wk.y * t2.y;
                    yy -= vk.x * t1.y + vk.y * t1.x + wk.x * t2.y + wk.y * t2.x;
                }
                const double ox = taur * yx - taui * yy, oy = taur * yy + taui * yx;
                const double2 v = vs[lr];
                dr += ox * v.x + oy * v.y;
                di += ox * v.y - oy * v.x;
                W[r + j * nn] = make_double2(ox, oy);
            }
            dr = block_sum<NT>(dr, red);
            di = block_sum<NT>(di, red);
            const double a2x = -0.5 * (taur * dr - taui * di), a2y = -0.5 * (taur * di + taui * dr);
            for (int r = s + tid; r < n; r += NT)
            {
                const double2 v = vs[r - s];
                double2 wr = W[r + j * nn];
                wr.x += a2x * v.x - a2y * v.y;
                wr.y += a2x * v.y + a2y * v.x;
                W[r + j * nn] = wr;
            }
            __syncthreads();
        }

        // trailing triangle: A(q0:n, q0:n) -= V W^H + W V^H (lower)
        const int q0 = p0 + EIG_NB;
        if (q0 < n)
        {
            const int ntt = (n - q0 + EIG_TILE - 1) / EIG_TILE;
            double2 *Vc = yp;
            double2 *Wc = yp + EIG_TILE * EIG_NB;
            for (int J = 0; J < ntt; ++J)
            {
                for (int t = tid; t < EIG_TILE * EIG_NB; t += NT)
                {
                    const int c = q0 + EIG_TILE * J + t / EIG_NB, k = t % EIG_NB;
                    Vc[t] = c < n ? A[c + (p0 + k) * nn] : zero;
                    Wc[t] = c < n ? W[c + k * nn] : zero;
                }
                __syncthreads();
                for (int I = J + wv; I < ntt; I += NW)
                {
                    const int r = q0 + EIG_TILE * I + rr;
                    const bool rv = r < n;
                    const int rs = min(r, n - 1);
                    const int c0 = q0 + EIG_TILE * J + 16 * h;
                    double ax[16], ay[16];
#pragma unroll
                    for (int cc = 0; cc < 16; ++cc)
                    {
                        const int c = c0 + cc;
                        const bool ok = rv && c < n && r >= c;
                        const double2 v = ok ? A[rs + c * nn] : zero;
                        ax[cc] = v.x;
                        ay[cc] = v.y;
                    }
                    for (int k = 0; k < EIG_NB; ++k)
                    {
                        const double2 vr = rv ? A[rs + (p0 + k) * nn] : zero, wr = rv ? W[rs + k * nn] : zero;
#pragma unroll
                        for (int cc = 0; cc < 16; ++cc)
                        {
                            const double2 vc = Vc[(16 * h + cc) * EIG_NB + k], wc = Wc[(16 * h + cc) * EIG_NB + k];
                            ax[cc] -= vr.x * wc.x + vr.y * wc.y + wr.x * vc.x + wr.y * vc.y;
                            ay[cc] -= vr.y * wc.x - vr.x * wc.y + wr.y * vc.x - wr.x * vc.y;
                        }
                    }
#pragma unroll
                    for (int cc = 0; cc < 16; ++cc)
                    {
                        const int c = c0 + cc;
                        if (rv && c < n && r >= c) A[r + c * nn] = make_double2(ax[cc], r == c ? 0.0 : ay[cc]);
                    }
                }
                __syncthreads();
            }
        }
    }
}

// Eigenvalues of the symmetric tridiagonal (d, e) by bisection on Sturm
// counts, and the von Neumann entropy in bits.
template <int NT>
__global__ void __launch_bounds__(NT) tridiag_entropy_kernel(const double *__restrict__ d, const double *__restrict__ e,
                                                             int n, double *__restrict__ lam, double *__restrict__ ent)
{
    constexpr int NW = NT / 64;
    extern __shared__ double sh[];
    double *sd = sh, *se2 = sh + n, *red = sh + 2 * n;  // red: NW
    const std::size_t off = blockIdx.x * static_cast<std::size_t>(n);
    const int tid = threadIdx.x;
    double lo = DBL_MAX, hi = -DBL_MAX, em = 0.0;
    for (int i = tid; i < n; i += NT)
    {
        const double di = d[off + i];
        const double el = i > 0 ? fabs(e[off + i - 1]) : 0.0, er = i + 1 < n ? fabs(e[off + i]) : 0.0;
        sd[i] = di;
        se2[i] = er * er;
        lo = fmin(lo, di - el - er);
        hi = fmax(hi, di + el + er);
        em = fmax(em, er * er);
    }
    // min / max over the workgroup (exact, order-free)
    for (int o = 32; o > 0; o >>= 1)
    {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
        em = fmax(em, __shfl_xor(em, o, 64));
    }
    __shared__ double mm[3][NW];
    if ((tid & 63) == 0)
    {
        mm[0][tid >> 6] = lo;
        mm[1][tid >> 6] = hi;
        mm[2][tid >> 6] = em;
    }
    __syncthreads();
    for (int w = 0; w < NW; ++w)
    {
        lo = fmin(lo, mm[0][w]);
        hi = fmax(hi, mm[1][w]);
        em = fmax(em, mm[2][w]);
    }
    const double pivmin = DBL_MIN * fmax(1.0, em);
    double s = 0.0;
    for (int k = tid; k < n; k += NT)
    {
        double a = lo, b = hi;
        for (int it = 0; it < 128; ++it)
        {
            const double m = 0.5 * (a + b);
            if (!(m > a && m < b)) break;
            if (b - a <= fmax(2.0 * DBL_EPSILON * fmax(fabs(a), fabs(b)), 1.0e-18)) break;
            double q = sd[0] - m;
            if (fabs(q) < pivmin) q = -pivmin;
            int c = q < 0.0;
            for (int i = 1; i < n; ++i)
            {
                q = (sd[i] - m) - se2[i - 1] / q;
                if (fabs(q) < pivmin) q = -pivmin;
                c += q < 0.0;
            }
            if (c > k) b = m;
            else a = m;
        }
        const double l = 0.5 * (a + b);
        if (lam) lam[off + k] = l;
        if (l > 1.0e-15) s += -l * log2(l);
    }
    s = block_sum<NT>(s, red);
    if (tid == 0) ent[blockIdx.x] = fabs(s) < 1.0e-12 ? 0.0 : s;
}

template <int NT>
int launch_pair(double2 *mats, int m, std::size_t stride, int count, double2 *wpanel, double *d, double *e, double *ent,
                double *lam, hipStream_t stream)
{
    constexpr int NW = NT / 64;
    const std::size_t lds = (static_cast<std::size_t>(m) + std::max<std::size_t>(static_cast<std::size_t>(NW) * m,
                                                                                 2 * EIG_TILE * EIG_NB) +
                             4 * EIG_NB) * sizeof(double2) + NW * sizeof(double);
    static std::size_t attr_bytes = 0;  // dynamic LDS limit raised so far (per instantiation)
    if (lds > attr_bytes)
    {
        MIPT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&hetrd_kernel<NT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        attr_bytes = lds;
    }
    HetrdArgs g{mats, stride, wpanel, static_cast<std::size_t>(m) * EIG_NB, d, e, m};
    hipLaunchKernelGGL(hetrd_kernel<NT>, dim3(static_cast<unsigned>(count)), dim3(NT), lds, stream, g);
    MIPT_HIP_CHECK(hipGetLastError());
    const std::size_t lds2 = (2 * static_cast<std::size_t>(m) + NW) * sizeof(double);
    hipLaunchKernelGGL(tridiag_entropy_kernel<NT>, dim3(static_cast<unsigned>(count)), dim3(NT), lds2, stream, d, e, m,
                       lam, ent);
    MIPT_HIP_CHECK(hipGetLastError());
    return OK;
}

} // namespace

int launch_hermitian_entropies(double2 *mats, int m, std::size_t stride, int count, double2 *wpanel, double *d,
                               double *e, double *ent, double *lam, hipStream_t stream)
{
    if (!mats || !wpanel || !d || !e || !ent || m < 1 || m > EIG_MAX_N || count < 0 || stride < static_cast<std::size_t>(m) * m)
    {
        set_error("hermitian entropies: bad arguments (m=%d, count=%d)", m, count);
        return E_ARG;
    }
    if (count == 0) return OK;
    if (m >= 512) return launch_pair<512>(mats, m, stride, count, wpanel, d, e, ent, lam, stream);
    if (m >= 64) return launch_pair<256>(mats, m, stride, count, wpanel, d, e, ent, lam, stream);
    return launch_pair<64>(mats, m, stride, count, wpanel, d, e, ent, lam, stream);
}

} // namespace mipt_hip
