// bd_matfree.hip -- matrix-free RPY products Y = D(r) X (no stored mobility).
//
// Replaces calc_rpy_tensor (bd_solver_m.f90:458-522) followed by the
// dsymv/dsymm('L') products (bd_solver_m.f90:446, brown_m.f90:232, 386)
// with one pass that evaluates every 3 x 3 bead-pair block of D in
// registers, where it is used, and never writes D:
//   D_ij X_j = c1 X_j + c2 r (r . X_j),   r = r_j - r_i,
//   r >= 2:  c1 = 3/(4r) + 1/(2r^3),   c2 = 3/(4r^3) - 3/(2r^5)
//   r <  2:  c1 = 1 - 9r/32,           c2 = 3/(32r)
//   i == j:  c1 = 1, c2 = 0
// (bd_solver_m.f90:484-504 with a = 1; c2 multiplies r r^T, not r^ r^T).
//
// Work decomposition.  A wavefront owns 64 row beads (R x 64 with R rows per
// lane); the NW waves of a workgroup share those rows and split the
// workgroup's range of column beads, and JG workgroups split the column
// range of a row block.  The workgroup first stages the records of its column
// beads -- coordinates (FP64) and the 3 x S block of X -- in LDS with vector
// loads all in flight at once, then every wave walks its share of them, all
// lanes on the same bead j, so each record is an LDS broadcast read whose
// latency the in-order LDS queue lets the loop hide.  The waves' sums are
// added in LDS in wave order and the JG workgroup sums by the next kernel in
// a fixed order, so results repeat bit for bit from run to run.
//
// Precision.  FP64: everything in FP64 (1/r from v_rsq_f64 refined by one
// third-order step).  FP32: the coordinate differences are formed in FP64 and
// rounded (so near pairs keep their relative accuracy however far the chain
// is from the origin), the block and its product in FP32, and the products
// are summed in FP32 over 8 column beads, then in FP64.
//
// Records are padded to a multiple of 8 beads with beads far away whose X
// block is zero, so the inner loop has no tail.
#include <mutex>
#include <string>

#include "bd_internal.h"
#include "bd_lanczos_dev.h"
#include "bd_matfree_dev.h"

namespace bd {

constexpr int MF_JG_MAX = 64;   // largest workgroup split of the column range

int mf_rows_per_lane(const Ctx &) { return 1; }

// Workgroups along the column range: about 4096 waves in total (sixteen
// per CU), at least one 8-bead chunk per wave.
static int mf_jg_for(int nb, int R)
{
    const int nib = (nb + 64 * R - 1) / (64 * R);
    const int nch = (nb + MF_CH - 1) / MF_CH;
    int jg = 4096 / (nib * MF_WAVES);
    jg = std::min(jg, nch / MF_WAVES);
    return std::max(1, std::min(jg, MF_JG_MAX));
}
int mf_jg(const Ctx &c) { return mf_jg_for(c.nb, mf_rows_per_lane(c)); }

// Coordinates (3, N) -> records of 4 doubles (x, y, z, 0), padded to nbp
// beads with points 1e6 apart from everything.
__global__ void mf_crd_kernel(const double *__restrict__ crd, int nb, int nbp, double *__restrict__ rec)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nbp) return;
    double x, y, z;
    if (j < nb) {
        x = crd[3 * j];
        y = crd[3 * j + 1];
        z = crd[3 * j + 2];
    } else {
        mf_far_point(j, nb, x, y, z);
    }
    rec[4 * j] = x;
    rec[4 * j + 1] = y;
    rec[4 * j + 2] = z;
    rec[4 * j + 3] = 0.0;
}

// X (rows x S column-major, ld) -> row-major records xr[r * S + k] in T,
// rows r < nrows (rows at or beyond n are zero).
template <typename T>
__global__ void mf_pack_kernel(const double *__restrict__ x, int ld, int n, int nrows, int S, T *__restrict__ xr)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nrows * S) return;
    const int r = (int)(e / S), k = (int)(e % S);
    xr[e] = r < n ? (T)x[(long)k * ld + r] : (T)0;
}

// part[g][i][a S + k] = sum over the column beads j of workgroup column g of
// (D_ij X_j)(a, k), FP64.  grid (ceil(nb / 64), JG), 64 MF_WAVES threads.
template <typename T, int S>
__global__ __launch_bounds__(64 * MF_WAVES) void mf_product_kernel(const double *__restrict__ crec,
                                                                   const T *__restrict__ xrec, int nb, int nch,
                                                                   double *__restrict__ part,
                                                                   const DevStatus *stop)
{
    using Rec = MfRec<T, S>;
    __shared__ __align__(16) unsigned char smem[mf_smem_bytes<T, S>()];
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = (stop && lz_done(stop)) ? 1 : 0;
    __syncthreads();
    if (s_done) return;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ic = min((int)blockIdx.x * 64 + lane, nb - 1);
    const double xi = crec[4 * ic], yi = crec[4 * ic + 1], zi = crec[4 * ic + 2];
    double acc[3][S];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int k = 0; k < S; k++) acc[a][k] = 0.0;
    // the workgroup's column chunks
    const int cb = (int)((long)nch * blockIdx.y / gridDim.y), ce = (int)((long)nch * (blockIdx.y + 1) / gridDim.y);
    auto stage = [&](int, int j, Rec &rc) {
        rc.x = crec[4 * j];
        rc.y = crec[4 * j + 1];
        rc.z = crec[4 * j + 2];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int k = 0; k < S; k++) rc.v[a][k] = xrec[(3 * j + a) * S + k];
    };
    mf_rows_times_cols<T, S>(xi, yi, zi, cb, ce, reinterpret_cast<Rec *>(smem), w, stage, acc);
    mf_wave_sums<S>(acc, reinterpret_cast<double(*)[64][3 * S + 1]>(smem), w, lane, [&](int l, int col, double v) {
        const int ib = (int)blockIdx.x * 64 + l;
        if (ib < nb) part[((long)blockIdx.y * nb + ib) * (3 * S) + col] = v;
    });
}

// y(rows, S) = sum over g of the partials (fixed order); rows >= n are zero.
__global__ void mf_reduce_kernel(const double *__restrict__ part, int jg, int nb, int n, int rows, int S,
                                 double *__restrict__ y, int ldy, const DevStatus *stop)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * S) return;
    if (stop && lz_done(stop)) return;
    const int row = (int)(e % rows), k = (int)(e / rows);
    double v = 0.0;
    if (row < n) {
        const long o = (long)(row / 3) * (3 * S) + (row % 3) * S + k, str = (long)nb * 3 * S;
        v = sum_strided(part + o, str, jg);
    }
    y[(long)k * ldy + row] = v;
}

// One 64-row block of a Lanczos iteration (grid npad / 64, 256 threads):
//   DX = sum of the partials (D applied to X, X = V_1, or Q from the second
//   iteration on);
//   with pre: R2 from the Gram partials of Q (every block, fixed order),
//     V_j = Q R2^{-1} (written), W = DX R2^{-1} (= D V_j by linearity),
//     B_{j-1} = R2 R1 (block 0 writes it);  else V_j = X, W = DX;
//   W1 = W - V_{j-1} B_{j-1}^T (stored over W);
//   partials of A = V_j^T W1 and P = W1^T W1 (2 S^2 per block, into lzpart),
// the same outputs as the stored-tile product's epilogue.
template <int S>
__global__ __launch_bounds__(256) void mf_reduce_lz_kernel(const double *__restrict__ part, int jg, int nb, int n,
                                                           const double *__restrict__ X, int ld, XPre pre,
                                                           const double *__restrict__ Vp, const double *__restrict__ Bp,
                                                           double *__restrict__ W, double *__restrict__ lzpart,
                                                           DevStatus *st)
{
    constexpr int NO = 2 * S * S, E = S * S;
    constexpr int EPT = (TILE * S + 255) / 256;   // elements of the block per thread
    constexpr int STR = 256 / E;                  // stripes over the Gram partial index
    constexpr int GPT = (GBLK + STR - 1) / STR;   // Gram partials per stripe, at most
    __shared__ double sdx[TILE][S + 1], sx[TILE][S + 1], sw[TILE][S + 1], sv[TILE][S + 1];
    __shared__ double sG[S * S], sR2[S * S], sB[S * S], sgs[256];
    __shared__ int s_done, s_fail;
    const int tid = threadIdx.x, G = blockIdx.x;
    const bool xpre = pre.gpart != nullptr;
    // every independent global load of the block first, all in flight
    // together (the status, this thread's stripe of the Gram partials, the
    // rows of X and V_{j-1}, then the column-group partials), and only then
    // the sums, in the same order as ever
    if (tid == 0) s_done = lz_done(st) ? 1 : 0;
    double gl[GPT];
    const int ge = tid % E, gs = tid / E;
#pragma unroll
    for (int q = 0; q < GPT; q++) {
        const int b = gs + q * STR;
        gl[q] = (xpre && b < pre.ng) ? pre.gpart[(long)b * E + ge] : 0.0;
    }
    double xv[EPT], vpv[EPT][S], dv[EPT];
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int e = tid + q * 256, r = e / S, k = e % S, row = G * TILE + r;
        const bool in = e < TILE * S;
        xv[q] = in ? X[(long)k * ld + row] : 0.0;
#pragma unroll
        for (int kp = 0; kp < S; kp++) vpv[q][kp] = (in && Vp) ? Vp[(long)kp * ld + row] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int e = tid + q * 256, r = e / S, k = e % S, row = G * TILE + r;
        double v = 0.0;
        if (e < TILE * S && row < n) {
            // all the partials of the element in flight at once (jg <= 64)
            const long o = (long)(row / 3) * (3 * S) + (row % 3) * S + k, str = (long)nb * 3 * S;
            v = sum_strided<false, 32>(part + o, str, jg);
        }
        dv[q] = v;
    }
    // the stripe sums of the Gram partials (partial index gs, gs + STR, ...)
    double gsum = 0.0;
#pragma unroll
    for (int q = 0; q < GPT; q++)
        if (gs + q * STR < pre.ng) gsum += gl[q];
    sgs[tid] = gsum;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int e = tid + q * 256, r = e / S, k = e % S;
        if (e < TILE * S) {
            sdx[r][k] = dv[q];
            sx[r][k] = xv[q];
        }
    }
    __syncthreads();
    if (s_done) return;
    if (xpre) {
        // the stripes in order, then R2 and B = R2 R1 (both upper triangular)
        if (tid < E) {
            double t = 0.0;
            for (int q = 0; q < STR; q++) t += sgs[q * E + tid];
            sG[tid] = t;
        }
        __syncthreads();
        if (tid == 0) {
            double R[S][S];
            const bool ok = small_chol<S>(sG, R);
            s_fail = ok ? 0 : 1;
            for (int i = 0; i < S; i++)
                for (int j = 0; j < S; j++) sR2[i + j * S] = R[i][j];
            if (ok) {
                for (int i = 0; i < S; i++)
                    for (int j = 0; j < S; j++) {
                        double t = 0.0;
                        for (int k = 0; k < S; k++) t += R[i][k] * pre.R1[k + j * S];
                        sB[i + j * S] = t;
                        if (G == 0) pre.Bout[i + j * S] = t;
                    }
            }
        }
        __syncthreads();
        if (s_fail) {
            if (G == 0 && tid == 0) lz_factor_failed(st, pre.blk);
            return;
        }
    } else if (Vp) {
        if (tid < S * S) sB[tid] = Bp[tid];
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int e = tid + q * 256, r = e / S, k = e % S, row = G * TILE + r;
        if (e >= TILE * S) break;
        double w, vj;
        if (xpre) {
            // forward solves of the row against R2, columns 0..k
            double tw[S], tv[S];
#pragma unroll
            for (int c = 0; c < S; c++) {
                if (c > k) break;
                double a = sdx[r][c], b = sx[r][c];
                for (int m = 0; m < c; m++) {
                    a -= tw[m] * sR2[m + c * S];
                    b -= tv[m] * sR2[m + c * S];
                }
                tw[c] = a / sR2[c + c * S];
                tv[c] = b / sR2[c + c * S];
            }
            w = tw[k];
            vj = tv[k];
            pre.V[(long)k * ld + row] = vj;
        } else {
            w = sdx[r][k];
            vj = sx[r][k];
        }
        if (Vp) {
            double t = 0.0;
#pragma unroll
            for (int kp = 0; kp < S; kp++) t += vpv[q][kp] * sB[k + kp * S];
            w = w - t;
        }
        W[(long)k * ld + row] = w;
        sw[r][k] = w;
        sv[r][k] = vj;
    }
    __syncthreads();
    // the 64-row column sums: RG row groups of 64 / RG rows per thread, then
    // the groups in order (fixed order)
    constexpr int RG = (NO <= 32) ? 8 : (NO <= 64 ? 4 : 2);
    __shared__ double sred[RG][NO];
    for (int e = tid; e < RG * NO; e += blockDim.x) {
        const int o = e % NO, gq = e / NO;
        const bool isA = o < S * S;
        const int oo = isA ? o : o - S * S, c = oo % S, k = oo / S;
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < TILE / RG; q++) {
            const int r = gq * (TILE / RG) + q;
            t += (isA ? sv[r][c] : sw[r][c]) * sw[r][k];
        }
        sred[gq][o] = t;
    }
    __syncthreads();
    for (int o = tid; o < NO; o += blockDim.x) {
        double t = sred[0][o];
#pragma unroll
        for (int gq = 1; gq < RG; gq++) t += sred[gq][o];
        lzpart[(long)G * NO + o] = t;
    }
}

// Dense n x n RPY matrix (column-major, unit diagonal) from the records,
// with the arithmetic of the tile assembly (calc_rpy_tensor).
__global__ void mf_dense_kernel(const double *__restrict__ crec, int nb, double *__restrict__ dense, int ld)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = 3L * nb;
    if (e >= n * n) return;
    const int col = (int)(e / n), row = (int)(e % n);
    const int bi = row / 3, bj = col / 3, da = row % 3, db = col % 3;
    double v;
    if (bi == bj) {
        v = da == db ? 1.0 : 0.0;
    } else {
        const double rx = crec[4 * bj] - crec[4 * bi], ry = crec[4 * bj + 1] - crec[4 * bi + 1],
                     rz = crec[4 * bj + 2] - crec[4 * bi + 2];
        const double rijm = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(rx, rx), __dmul_rn(ry, ry)), __dmul_rn(rz, rz)));
        const double irijm = 1.0 / rijm, irijm2 = irijm * irijm;
        double c1, c2, cons;
        if (rijm >= 2.0) {
            c1 = 1.0 + (2.0 / 3.0) * irijm2;
            c2 = 1.0 - 2.0 * irijm2;
            cons = 0.75 * irijm;
        } else {
            c1 = 1.0 - 9.0 * rijm / 32.0;
            c2 = 3.0 * rijm / 32.0;
            cons = 1.0;
        }
        const double ra = da == 0 ? rx : (da == 1 ? ry : rz), rb = db == 0 ? rx : (db == 1 ? ry : rz);
        v = cons * ((da == db ? c1 : 0.0) + c2 * ra * rb * irijm2);
    }
    dense[(long)col * ld + row] = v;
}

// ---------------------------------------------------------------- host
// BD_MOB_AUTO.  FP64: matrix-free from the bead count at which it was
// measured to beat the stored FP64 tiles.  FP32: stored tiles (FP32 storage,
// FP64 arithmetic), whose D.X error (~2e-8 relative, one rounding per entry)
// is a quarter of the FP32 matrix-free product's (~8e-8, several roundings
// per pair); the vector ALUs run FP32 and FP64 at one rate here (no packed
// FP32), so the FP32 matrix-free product is no faster than the FP64 one.
bool mob_use_matfree(const bd_opts &o, int nbeads)
{
    if (o.mob_storage == BD_MOB_TILES) return false;
    if (o.mob_storage == BD_MOB_MATFREE) return true;
    if (o.mob_fctr != BD_LANC) return false;   // the Cholesky path factors a stored matrix
    if (o.precision == BD_FP32) return false;
    return nbeads >= 64;
}

int mf_alloc(Ctx &c)
{
    c.mf_nbp = ((c.nb + MF_CH - 1) / MF_CH) * MF_CH;
    c.mf_rows = std::max(c.npad, 3 * c.mf_nbp);
    c.mf_jg = mf_jg(c);
    const size_t esz = c.opt.precision == BD_FP32 ? sizeof(float) : sizeof(double);
    BD_HIP(hipMalloc(&c.mf_crd, (size_t)c.mf_nbp * 4 * sizeof(double)));
    BD_HIP(hipMalloc(&c.mf_x, (size_t)c.mf_rows * MAX_S * esz));
    BD_HIP(hipMemset(c.mf_x, 0, (size_t)c.mf_rows * MAX_S * esz));
    BD_HIP(hipMalloc(&c.mf_part, (size_t)c.mf_jg * c.nb * 3 * MAX_S * sizeof(double)));
    return BD_OK;
}

void mf_free(Ctx &c)
{
    if (c.mf_crd) (void)hipFree(c.mf_crd);
    if (c.mf_x) (void)hipFree(c.mf_x);
    if (c.mf_part) (void)hipFree(c.mf_part);
    c.mf_crd = nullptr;
    c.mf_x = nullptr;
    c.mf_part = nullptr;
}

int mf_assemble(Ctx &c, const double *crd)
{
    hipLaunchKernelGGL(mf_crd_kernel, dim3((c.mf_nbp + 255) / 256), dim3(256), 0, c.stream, crd, c.nb, c.mf_nbp,
                       c.mf_crd);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int mf_pack(Ctx &c, const double *x, int s, int ldx)
{
    const long tot = (long)c.mf_rows * s;
    if (c.opt.precision == BD_FP32)
        hipLaunchKernelGGL(mf_pack_kernel<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, x, ldx,
                           c.n, c.mf_rows, s, (float *)c.mf_x);
    else
        hipLaunchKernelGGL(mf_pack_kernel<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, x,
                           ldx, c.n, c.mf_rows, s, (double *)c.mf_x);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

template <typename T, int S>
static int mf_launch_t(Ctx &c, const DevStatus *stop)
{
    const int R = mf_rows_per_lane(c);
    const int nch = c.mf_nbp / MF_CH;
    const dim3 grid((c.nb + 64 * R - 1) / (64 * R), c.mf_jg);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BD_TRY(prof_pair(c, S, &e0, &e1));
    if (e0) BD_HIP(hipEventRecord(e0, c.stream));
    const double *crec = c.mf_crd;
    const T *xrec = (const T *)c.mf_x;
    hipLaunchKernelGGL((mf_product_kernel<T, S>), grid, dim3(64 * MF_WAVES), 0, c.stream, crec, xrec, c.nb, nch,
                       c.mf_part, stop);
    BD_HIP(hipGetLastError());
    if (e1) BD_HIP(hipEventRecord(e1, c.stream));
    return BD_OK;
}

template <typename T>
static int mf_launch(Ctx &c, int s, const DevStatus *stop)
{
    switch (s) {
    case 1: return mf_launch_t<T, 1>(c, stop);
    case 2: return mf_launch_t<T, 2>(c, stop);
    case 3: return mf_launch_t<T, 3>(c, stop);
    case 4: return mf_launch_t<T, 4>(c, stop);
    case 5: return mf_launch_t<T, 5>(c, stop);
    case 6: return mf_launch_t<T, 6>(c, stop);
    case 7: return mf_launch_t<T, 7>(c, stop);
    case 8: return mf_launch_t<T, 8>(c, stop);
    default: set_error("block size s must be in 1..8"); return BD_ERR_ARG;
    }
}

// Product of the records already packed in c.mf_x (s columns) into the partials.
int mf_product_packed(Ctx &c, int s, const DevStatus *stop)
{
    if (c.opt.precision == BD_FP32) return mf_launch<float>(c, s, stop);
    return mf_launch<double>(c, s, stop);
}

int mf_product(Ctx &c, const double *x, double *y, int s, int ldx, int ldy, const DevStatus *stop)
{
    BD_TRY(mf_pack(c, x, s, ldx));
    BD_TRY(mf_product_packed(c, s, stop));
    const long tot = (long)c.npad * s;
    hipLaunchKernelGGL(mf_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, c.mf_part,
                       c.mf_jg, c.nb, c.n, c.npad, s, y, ldy, stop);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

template <int S>
static int mf_reduce_lz_t(Ctx &c, const double *X, double *W, int ld, const LzEpi &epi, const XPre *pre)
{
    const XPre xp = pre ? *pre : XPre{};
    hipLaunchKernelGGL((mf_reduce_lz_kernel<S>), dim3(c.npad / TILE), dim3(256), 0, c.stream, c.mf_part, c.mf_jg,
                       c.nb, c.n, X, ld, xp, epi.Vp, epi.Bp, W, epi.part, c.st);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int mf_product_lz(Ctx &c, const double *X, double *W, int s, int ld, bool packed, const LzEpi &epi,
                  const XPre *pre)
{
    if (!packed) BD_TRY(mf_pack(c, X, s, ld));
    BD_TRY(mf_product_packed(c, s, c.st));
    switch (s) {
    case 1: return mf_reduce_lz_t<1>(c, X, W, ld, epi, pre);
    case 2: return mf_reduce_lz_t<2>(c, X, W, ld, epi, pre);
    case 3: return mf_reduce_lz_t<3>(c, X, W, ld, epi, pre);
    case 4: return mf_reduce_lz_t<4>(c, X, W, ld, epi, pre);
    case 5: return mf_reduce_lz_t<5>(c, X, W, ld, epi, pre);
    case 6: return mf_reduce_lz_t<6>(c, X, W, ld, epi, pre);
    case 7: return mf_reduce_lz_t<7>(c, X, W, ld, epi, pre);
    case 8: return mf_reduce_lz_t<8>(c, X, W, ld, epi, pre);
    default: set_error("block size s must be in 1..8"); return BD_ERR_ARG;
    }
}

int mf_dense(Ctx &c, double *dense, int ld)
{
    const long tot = (long)c.n * c.n;
    if (tot == 0) return BD_OK;
    hipLaunchKernelGGL(mf_dense_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, c.mf_crd, c.nb,
                       dense, ld);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// ---------------------------------------------------------------- context-free product
// y(n, s) = D(crd) x(n, s) on device buffers with scratch scr (mf_raw_scratch
// doubles): the coordinate records, the packed block, then the partials.
// The column split jg is capped so that the partials fit MF_RAW_PART
// doubles, the same for every caller, so that results do not depend on
// where the scratch lives.
constexpr long MF_RAW_PART = 524288;   // 4 MiB of partial sums

static int mf_raw_jg(int nb, int s, int precision)
{
    const int jg = mf_jg_for(nb, 1);
    const long cap = MF_RAW_PART / ((long)nb * 3 * s);
    return (int)std::max(1L, std::min((long)jg, cap));
}

static long mf_raw_scratch(int nb, int s, int precision)
{
    const long nbp = (long)((nb + MF_CH - 1) / MF_CH) * MF_CH;
    const long esz = precision == BD_FP32 ? 4 : 8;
    return 4 * nbp + (3 * nbp * s * esz + 7) / 8 + (long)mf_raw_jg(nb, s, precision) * nb * 3 * s;
}

template <typename T, int S>
static void mf_raw_product(int nb, int nbp, int jg, const double *crec, const T *xrec, double *part,
                           hipStream_t st)
{
    const dim3 grid((nb + 63) / 64, jg);
    hipLaunchKernelGGL((mf_product_kernel<T, S>), grid, dim3(64 * MF_WAVES), 0, st, crec, xrec, nb, nbp / MF_CH,
                       part, (const DevStatus *)nullptr);
}

template <typename T>
static void mf_raw_product_s(int s, int nb, int nbp, int jg, const double *crec, const T *xrec, double *part,
                             hipStream_t st)
{
    switch (s) {
    case 1: mf_raw_product<T, 1>(nb, nbp, jg, crec, xrec, part, st); break;
    case 2: mf_raw_product<T, 2>(nb, nbp, jg, crec, xrec, part, st); break;
    case 3: mf_raw_product<T, 3>(nb, nbp, jg, crec, xrec, part, st); break;
    case 4: mf_raw_product<T, 4>(nb, nbp, jg, crec, xrec, part, st); break;
    case 5: mf_raw_product<T, 5>(nb, nbp, jg, crec, xrec, part, st); break;
    case 6: mf_raw_product<T, 6>(nb, nbp, jg, crec, xrec, part, st); break;
    case 7: mf_raw_product<T, 7>(nb, nbp, jg, crec, xrec, part, st); break;
    default: mf_raw_product<T, 8>(nb, nbp, jg, crec, xrec, part, st); break;
    }
}

static int mf_raw_check(int nbeads, const double *crd, const double *x, int ldx, int s, const double *y, int ldy,
                        int precision)
{
    if (nbeads < 1 || !crd || !x || !y || s < 1 || s > MAX_S || ldx < 3 * nbeads || ldy < 3 * nbeads ||
        (precision != BD_FP64 && precision != BD_FP32)) {
        set_error("bd_rpy_apply: bad arguments");
        return BD_ERR_ARG;
    }
    return BD_OK;
}

// Launches the whole product on stream st with scratch scr.
static int mf_raw_apply(int nb, const double *crd, const double *x, int ldx, int s, double *y, int ldy,
                        int precision, double *scr, hipStream_t st)
{
    const int nbp = ((nb + MF_CH - 1) / MF_CH) * MF_CH, rows = 3 * nbp, n = 3 * nb;
    const int jg = mf_raw_jg(nb, s, precision);
    const long esz = precision == BD_FP32 ? 4 : 8;
    double *crec = scr;
    void *xrec = scr + 4L * nbp;
    double *part = scr + 4L * nbp + (3L * nbp * s * esz + 7) / 8;
    hipLaunchKernelGGL(mf_crd_kernel, dim3((nbp + 255) / 256), dim3(256), 0, st, crd, nb, nbp, crec);
    const long tot = (long)rows * s;
    if (precision == BD_FP32) {
        hipLaunchKernelGGL(mf_pack_kernel<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, ldx, n,
                           rows, s, (float *)xrec);
        mf_raw_product_s<float>(s, nb, nbp, jg, crec, (const float *)xrec, part, st);
    } else {
        hipLaunchKernelGGL(mf_pack_kernel<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, ldx,
                           n, rows, s, (double *)xrec);
        mf_raw_product_s<double>(s, nb, nbp, jg, crec, (const double *)xrec, part, st);
    }
    const long ytot = (long)n * s;
    hipLaunchKernelGGL(mf_reduce_kernel, dim3((unsigned)((ytot + 255) / 256)), dim3(256), 0, st, part, jg, nb, n, n,
                       s, y, ldy, (const DevStatus *)nullptr);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// ---------------------------------------------------------------- no allocation
// Scratch of bd_rpy_apply_noalloc: one array per GPU, in the code object,
// for up to MF_NOALLOC_MAX_BEADS beads and s <= 8 (the records, the packed
// block, MF_RAW_PART doubles of partial sums): 753664 doubles, 5.75 MiB.
// Host code never uses the array's name as an address: the device address
// comes from hipGetSymbolAddress, once per device.
constexpr int MF_NOALLOC_MAX_BEADS = 8192;
constexpr long MF_NOALLOC_DOUBLES = 4L * MF_NOALLOC_MAX_BEADS + 3L * MF_NOALLOC_MAX_BEADS * MAX_S + MF_RAW_PART;
constexpr int MF_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[MF_NOALLOC_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MF_MAX_DEVICES] = {};

// The current device's address of scratch (resolved on its first use).
static double *noalloc_scratch(int dev)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("bd_rpy_apply_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace bd

using namespace bd;

extern "C" int bd_rpy_apply_noalloc(int nbeads, const double *crd, const double *x, int ldx, int s, double *y,
                                    int ldy, int precision, void *stream)
{
    BD_TRY(mf_raw_check(nbeads, crd, x, ldx, s, y, ldy, precision));
    if (nbeads > MF_NOALLOC_MAX_BEADS) {
        set_error("bd_rpy_apply_noalloc: nbeads exceeds " + std::to_string(MF_NOALLOC_MAX_BEADS) +
                  " (the size of its static scratch); use bd_rpy_apply");
        return BD_ERR_ARG;
    }
    if (mf_raw_scratch(nbeads, s, precision) > MF_NOALLOC_DOUBLES) {
        set_error("internal: bd_rpy_apply_noalloc scratch too small");
        return BD_ERR_STATE;
    }
    int dev = 0;
    BD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MF_MAX_DEVICES) { set_error("bd_rpy_apply_noalloc: device index out of range"); return BD_ERR_ARG; }
    double *scr = noalloc_scratch(dev);
    if (!scr) return BD_ERR_HIP;
    return mf_raw_apply(nbeads, crd, x, ldx, s, y, ldy, precision, scr, (hipStream_t)stream);
}

extern "C" int bd_rpy_apply(int nbeads, const double *crd, const double *x, int ldx, int s, double *y, int ldy,
                            int precision, void *stream)
{
    BD_TRY(mf_raw_check(nbeads, crd, x, ldx, s, y, ldy, precision));
    hipStream_t st = (hipStream_t)stream;
    double *scr = nullptr;
    BD_HIP(hipMalloc(&scr, (size_t)mf_raw_scratch(nbeads, s, precision) * sizeof(double)));
    int rc = mf_raw_apply(nbeads, crd, x, ldx, s, y, ldy, precision, scr, st);
    if (hipStreamSynchronize(st) != hipSuccess && rc == BD_OK) {
        set_error("bd_rpy_apply: stream synchronisation failed");
        rc = BD_ERR_HIP;
    }
    (void)hipFree(scr);
    return rc;
}
