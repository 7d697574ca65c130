// adjoint2d.hip -- per-cell-diffusivity ADI, its discrete adjoint and the
// gradient of a field misfit with respect to every cell's diffusivity and to
// the decay rate, with checkpoint-and-recompute.  No reference counterpart:
// the forward model generalises Diffusion2DADI (include/Hal/Pde/ADIequations.h:306-337)
// to a per-cell rate with the face convention of the per-cell explicit
// stencil (src/Pde/PDEequations.cpp:214-219; the face between cells i and i+1
// carries D_i), zero-flux boundaries, and appends MulAll(-k) + Update.
//
// Tick (N = X*Y cells, u in HAL's x*Y+y layout, uT its transpose):
//   s1 = (I + L_y/2) u;  s2 = (I - L_x/2)^{-1} s1;  s3 = (I + L_x/2) s2;
//   s4 = (I - L_y/2)^{-1} s3;  u' = s4 - k u.
// Every factor is symmetric, so the adjoint tick applies them transposed in
// reverse order:  m1 = (I - L_y/2)^{-1} lam;  m2 = (I - L_x/2)^{-1} (I + L_x/2) m1;
//   lam' = (I + L_y/2) m2 - k lam,
// and the per-tick gradient contributions are
//   dJ/dD_c -= 1/2 (dy m1 . dy s4 + dx m1 . dx s2 + dx m2 . dx s2 + dy m2 . dy u)
//   dJ/dk   -= <lam, u>
// with dx p = p(x+1,y) - p(x,y) (0 on the last row), dy likewise.
//
// Kernels: a line solve (one workgroup of four wavefronts per line,
// thread-local recurrences composed by wavefront scans of affine maps and
// LDS-exchanged wavefront totals; the per-cell Thomas factors are computed
// once per diffusivity field on the host in FP64), an LDS-tiled transpose so
// each sweep reads its lines contiguously, and a fused gradient /
// adjoint-update kernel.  Checkpoints every K ticks; the backward pass recomputes each
// segment from its checkpoint, keeping u, s2, s4 of the segment.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

using namespace halhip;

namespace
{
    template <int CTRL, int ROW_MASK = 0xF>
    __device__ __forceinline__ float DppF(float old, float v)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                      __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                                                      0xF, false));
    }

    struct LineArgs
    {
        int32_t nLines;
        int32_t L;
        const float *in;   // [nLines][L]
        const float *rate; // [nLines][L] face rates of this line family
        const float *fw;   // [nLines][L] Thomas factors
        const float *ib;
        const float *bw;
        float *out;        // [nLines][L]
        const float *prev; // decay: out = x - k * prev
        float negK;
        float *stash;      // optional copy of x before decay
    };

    // Line solve: a workgroup of kBlockLine threads per line (four
    // wavefronts, EPL adjacent elements per thread, L <= kBlockLine * EPL).
    // With VEC (EPL and L multiples of 4) every array moves as 16-byte
    // vectors.  Each recurrence composes per-thread folds, a wavefront scan
    // of affine maps (DPP) and the wavefront totals exchanged through LDS.
    constexpr int kBlockLine = 256;

    template <int EPL, bool VEC>
    __device__ __forceinline__ void LoadLine(const float *__restrict__ src, size_t base, int j0, int L, float pad,
                                             float (&v)[EPL])
    {
        if constexpr (VEC)
        {
            // L is a multiple of 4: each group of 4 lies wholly inside or
            // wholly past the line
#pragma unroll
            for (int q = 0; q < EPL / 4; q++)
            {
                float4 t = make_float4(pad, pad, pad, pad);
                if (j0 + 4 * q < L)
                {
                    t = *reinterpret_cast<const float4 *>(src + base + j0 + 4 * q);
                }
                v[4 * q] = t.x;
                v[4 * q + 1] = t.y;
                v[4 * q + 2] = t.z;
                v[4 * q + 3] = t.w;
            }
        }
        else
        {
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                v[e] = (j0 + e < L) ? src[base + j0 + e] : pad;
            }
        }
    }

    template <int EPL, bool VEC>
    __device__ __forceinline__ void StoreLine(float *__restrict__ dst, size_t base, int j0, int L, const float (&v)[EPL])
    {
        if constexpr (VEC)
        {
#pragma unroll
            for (int q = 0; q < EPL / 4; q++)
            {
                if (j0 + 4 * q < L)
                {
                    *reinterpret_cast<float4 *>(dst + base + j0 + 4 * q) =
                        make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
                }
            }
        }
        else
        {
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                if (j0 + e < L)
                {
                    dst[base + j0 + e] = v[e];
                }
            }
        }
    }

    // RHS 0: d = c + 1/2 (rate[l] (in[l+1] - c) + rate[l-1] (in[l-1] - c)) across lines.
    // RHS 1: the same along the line (neighbours j+-1 of the same line).
    // RHS 2: d = in.
    // DECAY: out = x - k * prev; STASH: keeps x before the decay.
    template <int EPL, int RHS, bool DECAY, bool STASH, bool VEC>
    __global__ __launch_bounds__(kBlockLine) void CellLineSolveBlock(LineArgs a)
    {
        __shared__ float sA[kBlockLine / 64];
        __shared__ float sB[kBlockLine / 64];
        const int t = threadIdx.x;
        const int lane = t & 63;
        const int wave = t >> 6;
        const int line = blockIdx.x;
        const int L = a.L;
        const size_t base = static_cast<size_t>(line) * L;
        const int j0 = t * EPL;
        float d[EPL];
        float fw[EPL];
        LoadLine<EPL, VEC>(a.in, base, j0, L, 0.0f, d);
        LoadLine<EPL, VEC>(a.fw, base, j0, L, 1.0f, fw);
        if constexpr (RHS == 0)
        {
            // explicit half-step across lines: neighbours in lines l +- 1
            float r[EPL];
            float nb[EPL];
            float s[EPL];
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                s[e] = 0.0f;
            }
            if (line + 1 < a.nLines)
            {
                LoadLine<EPL, VEC>(a.rate, base, j0, L, 0.0f, r);
                LoadLine<EPL, VEC>(a.in, base + L, j0, L, 0.0f, nb);
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    s[e] += r[e] * (nb[e] - d[e]);
                }
            }
            if (line > 0)
            {
                LoadLine<EPL, VEC>(a.rate, base - L, j0, L, 0.0f, r);
                LoadLine<EPL, VEC>(a.in, base - L, j0, L, 0.0f, nb);
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    s[e] += r[e] * (nb[e] - d[e]);
                }
            }
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                d[e] = (j0 + e < L) ? fmaf(0.5f, s[e], d[e]) : 0.0f;
            }
        }
        else if constexpr (RHS == 1)
        {
            // explicit half-step along the line; the outer neighbours of the
            // thread's elements come from the cache
            float r[EPL];
            LoadLine<EPL, VEC>(a.rate, base, j0, L, 0.0f, r);
            const float cLeft = (j0 > 0 && j0 - 1 < L) ? a.in[base + j0 - 1] : 0.0f;
            const float rLeft = (j0 > 0 && j0 - 1 < L) ? a.rate[base + j0 - 1] : 0.0f;
            const float cRight = (j0 + EPL < L) ? a.in[base + j0 + EPL] : 0.0f;
            float out[EPL];
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                const int j = j0 + e;
                const float c = d[e];
                const float cp = (e + 1 < EPL) ? d[e + 1] : cRight;
                const float cm = (e > 0) ? d[e - 1] : cLeft;
                const float rm = (e > 0) ? r[e - 1] : rLeft;
                float sum = 0.0f;
                sum += (j + 1 < L) ? r[e] * (cp - c) : 0.0f;
                sum += (j > 0) ? rm * (cm - c) : 0.0f;
                out[e] = (j < L) ? fmaf(0.5f, sum, c) : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                d[e] = out[e];
            }
        }

        // Forward recurrence d'_j = fw_j d'_{j-1} + d_j, carry into the line 0.
        float A = 1.0f;
        float B = 0.0f;
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            B = fmaf(fw[e], B, d[e]);
            A *= fw[e];
        }
        {
            float tb;
            tb = DppF<0x111>(0.0f, B);
            B = fmaf(A, tb, B);
            A *= DppF<0x111>(1.0f, A);
            tb = DppF<0x112>(0.0f, B);
            B = fmaf(A, tb, B);
            A *= DppF<0x112>(1.0f, A);
            tb = DppF<0x114>(0.0f, B);
            B = fmaf(A, tb, B);
            A *= DppF<0x114>(1.0f, A);
            tb = DppF<0x118>(0.0f, B);
            B = fmaf(A, tb, B);
            A *= DppF<0x118>(1.0f, A);
            tb = DppF<0x142, 0xA>(0.0f, B);
            B = fmaf(A, tb, B);
            A *= DppF<0x142, 0xA>(1.0f, A);
            tb = DppF<0x143, 0xC>(0.0f, B);
            B = fmaf(A, tb, B);
            A *= DppF<0x143, 0xC>(1.0f, A);
        }
        if (lane == 63)
        {
            sA[wave] = A;
            sB[wave] = B;
        }
        __syncthreads();
        float cw = 0.0f;
        for (int g = 0; g < wave; g++)
        {
            cw = fmaf(sA[g], cw, sB[g]);
        }
        // this lane's carry: the preceding lanes' composition applied to cw
        float acc = fmaf(DppF<0x138>(1.0f, A), cw, DppF<0x138>(0.0f, B));
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            acc = fmaf(fw[e], acc, d[e]);
            d[e] = acc;
        }

        // Backward recurrence x_j = bw_j x_{j+1} + ib_j d'_j from the right.
        float bw[EPL];
        float ib[EPL];
        LoadLine<EPL, VEC>(a.bw, base, j0, L, 0.0f, bw);
        LoadLine<EPL, VEC>(a.ib, base, j0, L, 0.0f, ib);
        A = 1.0f;
        B = 0.0f;
#pragma unroll
        for (int e = EPL - 1; e >= 0; e--)
        {
            B = fmaf(bw[e], B, ib[e] * d[e]);
            A *= bw[e];
        }
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int o = 1 << k;
            const float Bd = __shfl_down(B, o);
            const float Ad = __shfl_down(A, o);
            if (lane + o < 64)
            {
                B = fmaf(A, Bd, B);
                A *= Ad;
            }
        }
        __syncthreads(); // the forward totals are read
        if (lane == 0)
        {
            sA[wave] = A;
            sB[wave] = B;
        }
        __syncthreads();
        cw = 0.0f;
        for (int g = kBlockLine / 64 - 1; g > wave; g--)
        {
            cw = fmaf(sA[g], cw, sB[g]);
        }
        const float An = __shfl_down(A, 1);
        const float Bn = __shfl_down(B, 1);
        acc = (lane == 63) ? cw : fmaf(An, cw, Bn);
#pragma unroll
        for (int e = EPL - 1; e >= 0; e--)
        {
            acc = fmaf(bw[e], acc, ib[e] * d[e]);
            d[e] = acc;
        }
        if constexpr (STASH)
        {
            StoreLine<EPL, VEC>(a.stash, base, j0, L, d);
        }
        if constexpr (DECAY)
        {
            float pv[EPL];
            LoadLine<EPL, VEC>(a.prev, base, j0, L, 0.0f, pv);
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                d[e] = fmaf(a.negK, pv[e], d[e]);
            }
        }
        StoreLine<EPL, VEC>(a.out, base, j0, L, d);
    }

    // out[c][r] = in[r][c] for an R x C matrix (R, C multiples of 4) through
    // 32 x 32 LDS tiles, one 16-byte vector read and written per thread
    // (many small tiles keep every CU busy on a 1024^2 field).
    __global__ __launch_bounds__(256) void Transpose4(const float *__restrict__ in, float *__restrict__ out, int32_t R,
                                                      int32_t C)
    {
        __shared__ float tile[32][33];
        const int c0 = blockIdx.x * 32;
        const int r0 = blockIdx.y * 32;
        const int q = threadIdx.x & 7;   // 4-column group
        const int rr = threadIdx.x >> 3; // 0..31
        {
            const int r = r0 + rr;
            const int c = c0 + 4 * q;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (r < R && c < C)
            {
                v = *reinterpret_cast<const float4 *>(in + static_cast<size_t>(r) * C + c);
            }
            tile[rr][4 * q] = v.x;
            tile[rr][4 * q + 1] = v.y;
            tile[rr][4 * q + 2] = v.z;
            tile[rr][4 * q + 3] = v.w;
        }
        __syncthreads();
        {
            const int c = c0 + rr; // output row
            const int r = r0 + 4 * q;
            if (c < C && r < R)
            {
                *reinterpret_cast<float4 *>(out + static_cast<size_t>(c) * R + r) =
                    make_float4(tile[4 * q][rr], tile[4 * q + 1][rr], tile[4 * q + 2][rr], tile[4 * q + 3][rr]);
            }
        }
    }

    // out[c][r] = in[r][c] for an R x C matrix, through 64 x 64 LDS tiles.
    __global__ __launch_bounds__(256) void Transpose(const float *__restrict__ in, float *__restrict__ out, int32_t R,
                                                     int32_t C)
    {
        __shared__ float tile[64][65];
        const int c0 = blockIdx.x * 64;
        const int r0 = blockIdx.y * 64;
        const int tx = threadIdx.x & 63;
        const int ty = threadIdx.x >> 6;
        for (int k = ty; k < 64; k += 4)
        {
            const int r = r0 + k;
            const int c = c0 + tx;
            tile[k][tx] = (r < R && c < C) ? in[static_cast<size_t>(r) * C + c] : 0.0f;
        }
        __syncthreads();
        for (int k = ty; k < 64; k += 4)
        {
            const int c = c0 + k;
            const int r = r0 + tx;
            if (c < C && r < R)
            {
                out[static_cast<size_t>(c) * R + r] = tile[tx][k];
            }
        }
    }

    struct GradArgs
    {
        int32_t X;
        int32_t Y;
        const float *D;   // [X][Y]
        const float *m1;  // [X][Y]
        const float *m2;  // [X][Y]
        const float *s2;  // [X][Y]
        const float *s4;  // [X][Y]
        const float *u;   // [X][Y]
        const float *lam; // [X][Y]
        float *lamNext;   // [X][Y]
        float *gradD;     // [X][Y], accumulated
        double *partialK; // one per block
        float negK;
    };

    // Per-tick gradient contributions and lam' = (I + L_y/2) m2 - k lam.  The
    // decay gradient -<lam, u> cancels heavily (lam has near-zero mean), so
    // its products and sums are formed in FP64.
    __global__ __launch_bounds__(256) void CellGradient(GradArgs a)
    {
        __shared__ double red[4];
        const int y = blockIdx.x * 64 + (threadIdx.x & 63);
        const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
        double kPart = 0.0;
        if (x < a.X && y < a.Y)
        {
            const size_t i = static_cast<size_t>(x) * a.Y + y;
            const bool hx = x + 1 < a.X;
            const bool hy = y + 1 < a.Y;
            const size_t ix = hx ? i + a.Y : i;
            const size_t iy = hy ? i + 1 : i;
            const float dyM1 = a.m1[iy] - a.m1[i];
            const float dyS4 = a.s4[iy] - a.s4[i];
            const float dxM1 = a.m1[ix] - a.m1[i];
            const float dxS2 = a.s2[ix] - a.s2[i];
            const float dxM2 = a.m2[ix] - a.m2[i];
            const float dyM2 = a.m2[iy] - a.m2[i];
            const float dyU = a.u[iy] - a.u[i];
            const float g = dyM1 * dyS4 + (dxM1 + dxM2) * dxS2 + dyM2 * dyU;
            a.gradD[i] = fmaf(-0.5f, g, a.gradD[i]);
            const float lam = a.lam[i];
            kPart = -static_cast<double>(lam) * static_cast<double>(a.u[i]);
            // (I + L_y/2) m2 along the row
            const float c = a.m2[i];
            float s = 0.0f;
            if (hy)
            {
                s += a.D[i] * (a.m2[i + 1] - c);
            }
            if (y > 0)
            {
                s += a.D[i - 1] * (a.m2[i - 1] - c);
            }
            a.lamNext[i] = fmaf(a.negK, lam, fmaf(0.5f, s, c));
        }
        // Block sum of the decay-gradient part (deterministic: fixed tree).
        for (int o = 32; o > 0; o >>= 1)
        {
            kPart += __shfl_down(kPart, o);
        }
        if ((threadIdx.x & 63) == 0)
        {
            red[threadIdx.x >> 6] = kPart;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            a.partialK[blockIdx.y * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        }
    }

    // loss partials: 1/2 (u - target)^2 per block, and lam = u - target.
    __global__ __launch_bounds__(256) void LossSeed(const float *__restrict__ u, const float *__restrict__ target,
                                                    float *__restrict__ lam, int64_t n, double *__restrict__ partial)
    {
        __shared__ double red[4];
        double s = 0.0;
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            const float r = u[i] - target[i];
            lam[i] = r;
            s += 0.5 * static_cast<double>(r) * static_cast<double>(r);
        }
        for (int o = 32; o > 0; o >>= 1)
        {
            s += __shfl_down(s, o);
        }
        if ((threadIdx.x & 63) == 0)
        {
            red[threadIdx.x >> 6] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        }
    }

    // Thomas factors of the line family: lines of length L with face rates
    // rate[line][j] (face j..j+1), zero flux.
    void FactorFamily(int32_t nLines, int32_t L, const std::vector<double> &rate, std::vector<float> &fw,
                      std::vector<float> &ib, std::vector<float> &bw)
    {
        const size_t n = static_cast<size_t>(nLines) * L;
        fw.assign(n, 0.0f);
        ib.assign(n, 0.0f);
        bw.assign(n, 0.0f);
        for (int32_t l = 0; l < nLines; l++)
        {
            const size_t base = static_cast<size_t>(l) * L;
            double bpPrev = 0.0;
            for (int32_t j = 0; j < L; j++)
            {
                const double dm = (j > 0) ? rate[base + j - 1] : 0.0;
                const double dp = (j < L - 1) ? rate[base + j] : 0.0;
                const double aj = -0.5 * dm;
                const double cj = -0.5 * dp;
                const double cPrev = (j > 0) ? -0.5 * rate[base + j - 1] : 0.0;
                double bp = 1.0 + 0.5 * (dm + dp);
                double w = 0.0;
                if (j > 0)
                {
                    w = aj / bpPrev;
                    bp -= w * cPrev;
                }
                fw[base + j] = static_cast<float>(-w);
                ib[base + j] = static_cast<float>(1.0 / bp);
                bw[base + j] = static_cast<float>((j < L - 1) ? -cj / bp : 0.0);
                bpPrev = bp;
            }
        }
    }

}

struct halhip_adjoint2d
{
    halhip_context *ctx = nullptr;
    int32_t X = 0;
    int32_t Y = 0;
    int32_t ticks = 0;
    int32_t every = 1;
    float decay = 0.0f;
    bool ready = false;
    DeviceArray<float> D, DT;              // rates in both layouts
    DeviceArray<float> fA, iA, bA;         // x-line factors, [Y][X]
    DeviceArray<float> fB, iB, bB;         // y-line factors, [X][Y]
    DeviceArray<float> u, uT, s2T, work;   // field work buffers
    DeviceArray<float> ckpt;               // (ticks/every + 1) checkpoints
    DeviceArray<float> stashU, stashS2, stashS4; // one segment
    DeviceArray<float> lam, lamNext, m1, m1T, m2T, m2, grad;
    DeviceArray<double> partialK;
    DeviceArray<double> partialJ;
    StreamClock clock;
    Profiler prof;
};

namespace
{
    size_t Cells(const halhip_adjoint2d *p)
    {
        return static_cast<size_t>(p->X) * p->Y;
    }

    void RequireAdj(const halhip_adjoint2d *p, const char *caller)
    {
        Require(p != nullptr, std::string(caller) + ": null plan");
        Bind(p->ctx);
    }

    template <int EPL, int RHS, bool DECAY, bool STASH>
    void LaunchBlockLine(halhip_adjoint2d *p, const LineArgs &a)
    {
        const bool vec = (EPL % 4 == 0) && (a.L % 4 == 0);
        if (vec)
        {
            hipLaunchKernelGGL((CellLineSolveBlock<EPL, RHS, DECAY, STASH, (EPL % 4 == 0)>),
                               dim3(static_cast<unsigned>(a.nLines)), dim3(kBlockLine), 0, p->clock.stream, a);
        }
        else
        {
            hipLaunchKernelGGL((CellLineSolveBlock<EPL, RHS, DECAY, STASH, false>),
                               dim3(static_cast<unsigned>(a.nLines)), dim3(kBlockLine), 0, p->clock.stream, a);
        }
    }

    // One line solve per line of the family: a workgroup of four
    // wavefronts per line, EPL = the smallest power of two with
    // 256 * EPL >= L (sides are at most 4096).
    template <int RHS, bool DECAY, bool STASH>
    void LaunchLine(halhip_adjoint2d *p, const LineArgs &a, const char *name)
    {
        p->prof.Begin(p->clock.stream, name);
        int epl = 1;
        while (kBlockLine * epl < a.L)
        {
            epl *= 2;
        }
        switch (epl)
        {
            case 1:
                LaunchBlockLine<1, RHS, DECAY, STASH>(p, a);
                break;
            case 2:
                LaunchBlockLine<2, RHS, DECAY, STASH>(p, a);
                break;
            case 4:
                LaunchBlockLine<4, RHS, DECAY, STASH>(p, a);
                break;
            case 8:
                LaunchBlockLine<8, RHS, DECAY, STASH>(p, a);
                break;
            default:
                LaunchBlockLine<16, RHS, DECAY, STASH>(p, a);
                break;
        }
        Check(hipGetLastError(), name);
        p->prof.End(p->clock.stream);
    }

    void LaunchTranspose(halhip_adjoint2d *p, const float *in, float *out, int32_t R, int32_t C)
    {
        p->prof.Begin(p->clock.stream, "adjoint_transpose");
        if (R % 4 == 0 && C % 4 == 0)
        {
            const dim3 grid(static_cast<unsigned>((C + 31) / 32), static_cast<unsigned>((R + 31) / 32));
            hipLaunchKernelGGL(Transpose4, grid, dim3(256), 0, p->clock.stream, in, out, R, C);
        }
        else
        {
            const dim3 grid(static_cast<unsigned>((C + 63) / 64), static_cast<unsigned>((R + 63) / 64));
            hipLaunchKernelGGL(Transpose, grid, dim3(256), 0, p->clock.stream, in, out, R, C);
        }
        Check(hipGetLastError(), "Transpose launch");
        p->prof.End(p->clock.stream);
    }

    LineArgs FamilyA(halhip_adjoint2d *p)
    {
        LineArgs a{};
        a.nLines = p->Y;
        a.L = p->X;
        a.rate = p->DT.get();
        a.fw = p->fA.get();
        a.ib = p->iA.get();
        a.bw = p->bA.get();
        return a;
    }

    LineArgs FamilyB(halhip_adjoint2d *p)
    {
        LineArgs a{};
        a.nLines = p->X;
        a.L = p->Y;
        a.rate = p->D.get();
        a.fw = p->fB.get();
        a.ib = p->iB.get();
        a.bw = p->bB.get();
        return a;
    }

    // One forward tick from `in` to `out` (both [X][Y]); with stash pointers,
    // keeps s2 and s4 ([X][Y]).
    void ForwardTick(halhip_adjoint2d *p, const float *in, float *out, float *s2, float *s4)
    {
        const int32_t X = p->X;
        const int32_t Y = p->Y;
        LaunchTranspose(p, in, p->uT.get(), X, Y);
        LineArgs a = FamilyA(p);
        a.in = p->uT.get();
        a.out = p->s2T.get();
        LaunchLine<0, false, false>(p, a, "adjoint_fwd_solve_x");
        float *s2dst = (s2 != nullptr) ? s2 : p->work.get();
        LaunchTranspose(p, p->s2T.get(), s2dst, Y, X);
        LineArgs b = FamilyB(p);
        b.in = s2dst;
        b.out = out;
        b.prev = in;
        b.negK = -p->decay;
        if (s4 != nullptr)
        {
            b.stash = s4;
            LaunchLine<0, true, true>(p, b, "adjoint_fwd_solve_y");
        }
        else
        {
            LaunchLine<0, true, false>(p, b, "adjoint_fwd_solve_y");
        }
    }

    // Runs the forward model from u0 (device) storing checkpoints; result in p->u.
    void Forward(halhip_adjoint2d *p)
    {
        const size_t n = Cells(p);
        Check(hipMemcpyAsync(p->ckpt.get(), p->u.get(), n * sizeof(float), hipMemcpyDeviceToDevice, p->clock.stream),
              "hipMemcpyAsync");
        for (int32_t t = 0; t < p->ticks; t++)
        {
            ForwardTick(p, p->u.get(), p->lamNext.get(), nullptr, nullptr);
            p->u.SwapWith(p->lamNext);
            if ((t + 1) % p->every == 0 && t + 1 < p->ticks)
            {
                const size_t slot = static_cast<size_t>((t + 1) / p->every);
                Check(hipMemcpyAsync(p->ckpt.get() + slot * n, p->u.get(), n * sizeof(float), hipMemcpyDeviceToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
            }
        }
    }
}

extern "C"
{
    int halhip_adjoint2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_adjoint2d **out)
    {
        halhip_adjoint2d *p = nullptr;
        int rc = Guard(
            [&]
            {
                Require(out != nullptr, "halhip_adjoint2d_create: null out");
                *out = nullptr;
                Require(ctx != nullptr, "halhip_adjoint2d_create: null context");
                Require(xDim > 0 && yDim > 0, "halhip_adjoint2d_create: dimensions must be positive");
                if (xDim > 4096 || yDim > 4096)
                {
                    Fail(HALHIP_ERR_UNSUPPORTED, "halhip_adjoint2d_create: sides are limited to 4096");
                }
                Bind(ctx);
                p = new halhip_adjoint2d();
                p->ctx = ctx;
                p->X = xDim;
                p->Y = yDim;
                const size_t n = Cells(p);
                for (DeviceArray<float> *b : {&p->D, &p->DT, &p->fA, &p->iA, &p->bA, &p->fB, &p->iB, &p->bB, &p->u,
                                              &p->uT, &p->s2T, &p->work, &p->lam, &p->lamNext, &p->m1, &p->m1T,
                                              &p->m2T, &p->m2, &p->grad})
                {
                    b->Allocate(ctx, n);
                }
                p->partialJ.Allocate(ctx, 1024);
                p->clock.Create();
                *out = p;
            });
        if (rc != HALHIP_OK && p != nullptr)
        {
            p->clock.Destroy();
            delete p;
        }
        return rc;
    }

    int halhip_adjoint2d_destroy(halhip_adjoint2d *p)
    {
        return Guard(
            [&]
            {
                if (p == nullptr)
                {
                    return;
                }
                Bind(p->ctx);
                p->clock.Destroy();
                delete p;
            });
    }

    int halhip_adjoint2d_set_problem(halhip_adjoint2d *p, const double *diffusivity, double decay, int32_t ticks,
                                     int32_t checkpointEvery)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_set_problem");
                Require(diffusivity != nullptr, "halhip_adjoint2d_set_problem: null diffusivity");
                Require(ticks >= 1 && checkpointEvery >= 1, "halhip_adjoint2d_set_problem: ticks and checkpoint >= 1");
                Require(std::isfinite(decay), "halhip_adjoint2d_set_problem: bad decay");
                const int32_t X = p->X;
                const int32_t Y = p->Y;
                const size_t n = Cells(p);
                std::vector<double> Dxy(diffusivity, diffusivity + n);
                std::vector<double> Dyx(n);
                for (int32_t x = 0; x < X; x++)
                {
                    for (int32_t y = 0; y < Y; y++)
                    {
                        const double v = Dxy[static_cast<size_t>(x) * Y + y];
                        Require(std::isfinite(v) && v >= 0.0, "halhip_adjoint2d_set_problem: diffusivity must be finite "
                                                               "and non-negative");
                        Dyx[static_cast<size_t>(y) * X + x] = v;
                    }
                }
                std::vector<float> f, i, b;
                auto up = [&](DeviceArray<float> &dst, const std::vector<float> &src)
                {
                    Check(hipMemcpy(dst.get(), src.data(), src.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
                };
                std::vector<float> Df(Dxy.begin(), Dxy.end()), DTf(Dyx.begin(), Dyx.end());
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                up(p->D, Df);
                up(p->DT, DTf);
                FactorFamily(Y, X, Dyx, f, i, b);
                up(p->fA, f);
                up(p->iA, i);
                up(p->bA, b);
                FactorFamily(X, Y, Dxy, f, i, b);
                up(p->fB, f);
                up(p->iB, i);
                up(p->bB, b);
                p->decay = static_cast<float>(decay);
                p->ticks = ticks;
                p->every = checkpointEvery;
                const size_t nCk = static_cast<size_t>((ticks + checkpointEvery - 1) / checkpointEvery);
                p->ckpt.Allocate(p->ctx, nCk * n);
                const size_t seg = static_cast<size_t>(checkpointEvery);
                p->stashU.Allocate(p->ctx, seg * n);
                p->stashS2.Allocate(p->ctx, seg * n);
                p->stashS4.Allocate(p->ctx, seg * n);
                const size_t gblocks = static_cast<size_t>((Y + 63) / 64) * static_cast<size_t>((X + 3) / 4);
                p->partialK.Allocate(p->ctx, gblocks * static_cast<size_t>(ticks));
                p->ready = true;
            });
    }

    int halhip_adjoint2d_forward(halhip_adjoint2d *p, const float *u0, float *uT)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_forward");
                Require(p->ready, "halhip_adjoint2d_forward: call halhip_adjoint2d_set_problem first");
                Require(u0 != nullptr && uT != nullptr, "halhip_adjoint2d_forward: null array");
                const size_t n = Cells(p);
                Check(hipMemcpyAsync(p->u.get(), u0, n * sizeof(float), hipMemcpyHostToDevice, p->clock.stream),
                      "hipMemcpyAsync");
                Forward(p);
                Check(hipMemcpyAsync(uT, p->u.get(), n * sizeof(float), hipMemcpyDeviceToHost, p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->prof.Collect();
            });
    }

    int halhip_adjoint2d_gradient(halhip_adjoint2d *p, const float *u0, const float *target, double *loss,
                                  float *gradDiffusivity, double *gradDecay)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_gradient");
                Require(p->ready, "halhip_adjoint2d_gradient: call halhip_adjoint2d_set_problem first");
                Require(u0 != nullptr && target != nullptr && loss != nullptr && gradDiffusivity != nullptr &&
                            gradDecay != nullptr,
                        "halhip_adjoint2d_gradient: null array");
                const int32_t X = p->X;
                const int32_t Y = p->Y;
                const size_t n = Cells(p);
                hipStream_t s = p->clock.stream;
                Check(hipMemcpyAsync(p->u.get(), u0, n * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
                Forward(p);
                // Seed: lam = u_T - target, J = 1/2 |lam|^2.  The target goes
                // through the work buffer.
                Check(hipMemcpyAsync(p->work.get(), target, n * sizeof(float), hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
                const unsigned lossBlocks = 1024;
                hipLaunchKernelGGL(LossSeed, dim3(lossBlocks), dim3(256), 0, s, p->u.get(), p->work.get(), p->lam.get(),
                                   static_cast<int64_t>(n), p->partialJ.get());
                Check(hipGetLastError(), "LossSeed launch");
                Check(hipMemsetAsync(p->grad.get(), 0, n * sizeof(float), s), "hipMemsetAsync");
                const dim3 gGrid(static_cast<unsigned>((Y + 63) / 64), static_cast<unsigned>((X + 3) / 4));
                const size_t gBlocks = static_cast<size_t>(gGrid.x) * gGrid.y;
                std::vector<double> hostK(gBlocks * static_cast<size_t>(p->ticks));
                double gk = 0.0;
                const int32_t nSeg = (p->ticks + p->every - 1) / p->every;
                for (int32_t seg = nSeg - 1; seg >= 0; seg--)
                {
                    const int32_t t0 = seg * p->every;
                    const int32_t t1 = std::min(p->ticks, t0 + p->every);
                    // Recompute the segment from its checkpoint, keeping u, s2, s4.
                    Check(hipMemcpyAsync(p->stashU.get(), p->ckpt.get() + static_cast<size_t>(seg) * n,
                                         n * sizeof(float), hipMemcpyDeviceToDevice, s),
                          "hipMemcpyAsync");
                    for (int32_t t = t0; t < t1; t++)
                    {
                        const size_t k = static_cast<size_t>(t - t0);
                        float *next = (t + 1 < t1) ? p->stashU.get() + (k + 1) * n : p->work.get();
                        ForwardTick(p, p->stashU.get() + k * n, next, p->stashS2.get() + k * n,
                                    p->stashS4.get() + k * n);
                    }
                    // Adjoint ticks backwards through the segment.
                    for (int32_t t = t1 - 1; t >= t0; t--)
                    {
                        const size_t k = static_cast<size_t>(t - t0);
                        LineArgs b = FamilyB(p);
                        b.in = p->lam.get();
                        b.out = p->m1.get();
                        LaunchLine<2, false, false>(p, b, "adjoint_solve_y");
                        LaunchTranspose(p, p->m1.get(), p->m1T.get(), X, Y);
                        LineArgs a = FamilyA(p);
                        a.in = p->m1T.get();
                        a.out = p->m2T.get();
                        LaunchLine<1, false, false>(p, a, "adjoint_solve_x");
                        LaunchTranspose(p, p->m2T.get(), p->m2.get(), Y, X);
                        GradArgs g{};
                        g.X = X;
                        g.Y = Y;
                        g.D = p->D.get();
                        g.m1 = p->m1.get();
                        g.m2 = p->m2.get();
                        g.s2 = p->stashS2.get() + k * n;
                        g.s4 = p->stashS4.get() + k * n;
                        g.u = p->stashU.get() + k * n;
                        g.lam = p->lam.get();
                        g.lamNext = p->lamNext.get();
                        g.gradD = p->grad.get();
                        g.partialK = p->partialK.get() + static_cast<size_t>(t) * gBlocks;
                        g.negK = -p->decay;
                        p->prof.Begin(s, "adjoint_gradient");
                        hipLaunchKernelGGL(CellGradient, gGrid, dim3(256), 0, s, g);
                        Check(hipGetLastError(), "CellGradient launch");
                        p->prof.End(s);
                        p->lam.SwapWith(p->lamNext);
                    }
                }
                std::vector<double> hostJ(lossBlocks);
                Check(hipMemcpyAsync(hostJ.data(), p->partialJ.get(), lossBlocks * sizeof(double),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
                Check(hipMemcpyAsync(gradDiffusivity, p->grad.get(), n * sizeof(float), hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
                Check(hipMemcpyAsync(hostK.data(), p->partialK.get(), hostK.size() * sizeof(double),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(s), "hipStreamSynchronize");
                double J = 0.0;
                for (double v : hostJ)
                {
                    J += v;
                }
                for (double v : hostK)
                {
                    gk += v;
                }
                *loss = J;
                *gradDecay = gk;
                p->prof.Collect();
            });
    }

    int halhip_adjoint2d_set_profiling(halhip_adjoint2d *p, int32_t on)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_set_profiling");
                p->prof.Reset();
                p->prof.SetOn(on != 0);
            });
    }

    int halhip_adjoint2d_profile(halhip_adjoint2d *p, int32_t maxKernels, int32_t *nKernels, char *names,
                                 int32_t namesLen, double *totalMs, int64_t *launches)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_profile");
                p->prof.Export(maxKernels, nKernels, names, namesLen, totalMs, launches);
            });
    }
}
