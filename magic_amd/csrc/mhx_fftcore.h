// Building blocks shared by the FFT kernels (mhx_fft.hip, mhx_grid.hip, the
// batched jobs in mhx_nlsplit.hip): complex helpers, in-register DFTs of
// sizes 2, 3, 4, 5, 12 and 16 with the kernel exp(sign 2 pi i nk / R), the
// wave-local LDS ordering point, the per-level Courant maxima of the grid
// stages, and generic Stockham row FFTs in LDS with the real-FFT pre/post
// steps.
#pragma once

#include "mhx_internal.h"

namespace mhx {

namespace {

template <typename T> struct CT;
template <> struct CT<double> { typedef double2 t; };
template <> struct CT<float> { typedef float2 t; };
template <typename T> using cpx = typename CT<T>::t;

template <typename C> __device__ inline C mk(decltype(C::x) a, decltype(C::x) b)
{
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C> __device__ inline C cmul(C a, C b)
{
    return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ inline C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C> __device__ inline C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ inline C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ inline C cmuli(C a, int sign)  // * (sign i)
{
    return sign > 0 ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x);
}

// in-register DFT of size R with kernel exp(sign 2 pi i / R)
template <int R, typename C> struct Dft;
template <typename C> struct Dft<2, C> {
    __device__ static void run(C *x, int)
    {
        C a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};
template <typename C> struct Dft<4, C> {
    __device__ static void run(C *x, int sign)
    {
        C a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
        C b0 = cadd(x[1], x[3]), b1 = cmuli(csub(x[1], x[3]), sign);
        x[0] = cadd(a0, b0);
        x[2] = csub(a0, b0);
        x[1] = cadd(a1, b1);
        x[3] = csub(a1, b1);
    }
};
template <typename C> struct Dft<3, C> {
    __device__ static void run(C *x, int sign)
    {
        typedef decltype(C::x) T;
        const T c = T(-0.5), s = T(0.86602540378443864676) * T(sign);
        C t1 = cadd(x[1], x[2]), t2 = csub(x[1], x[2]);
        C y0 = cadd(x[0], t1);
        C m1 = mk<C>(x[0].x + c * t1.x, x[0].y + c * t1.y);
        C m2 = mk<C>(-s * t2.y, s * t2.x);
        x[0] = y0;
        x[1] = cadd(m1, m2);
        x[2] = csub(m1, m2);
    }
};
template <typename C> struct Dft<5, C> {
    __device__ static void run(C *x, int sign)
    {
        typedef decltype(C::x) T;
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212) * T(sign), s2 = T(0.58778525229247312917) * T(sign);
        C a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
        C a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
        C y0 = cadd(x[0], cadd(a1, a2));
        C r1 = mk<C>(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
        C r2 = mk<C>(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
        C i1 = mk<C>(-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x);
        C i2 = mk<C>(-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x);
        x[0] = y0;
        x[1] = cadd(r1, i1);
        x[4] = csub(r1, i1);
        x[2] = cadd(r2, i2);
        x[3] = csub(r2, i2);
    }
};

// 12-point DFT as 4 x 3 with the prime-factor (Good-Thomas) index maps
// n = 3 n1 + 4 n2, k = 9 k1 + 4 k2 (mod 12): no internal twiddles
template <typename C> struct Dft<12, C> {
    __device__ __forceinline__ static void run(C *x, int sign)
    {
        C y[4][3];
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2) {
            C a[4];
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) a[n1] = x[(3 * n1 + 4 * n2) % 12];
            Dft<4, C>::run(a, sign);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) y[k1][n2] = a[k1];
        }
        C out[12];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            C b[3] = {y[k1][0], y[k1][1], y[k1][2]};
            Dft<3, C>::run(b, sign);
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) out[(9 * k1 + 4 * k2) % 12] = b[k2];
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) x[i] = out[i];
    }
};

// 16-point DFT as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2; inner DFTs over n1,
// twiddles exp(sign 2 pi i n2 k1 / 16), outer DFTs over n2
template <typename C> struct Dft<16, C> {
    __device__ __forceinline__ static void run(C *x, int sign)
    {
        typedef decltype(C::x) T;
        const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);
        const T r2 = T(0.70710678118654752440), sg = T(sign);
        C a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            C t[4] = {x[n2], x[4 + n2], x[8 + n2], x[12 + n2]};
            Dft<4, C>::run(t, sign);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
        }
        // exp(sign 2 pi i e / 16) for e = n2 k1
        auto tw = [&](C v, int e) -> C {
            switch (e) {
            case 1: return mk<C>(c1 * v.x - sg * s1 * v.y, c1 * v.y + sg * s1 * v.x);
            case 2: return mk<C>(r2 * (v.x - sg * v.y), r2 * (v.y + sg * v.x));
            case 3: return mk<C>(s1 * v.x - sg * c1 * v.y, s1 * v.y + sg * c1 * v.x);
            case 4: return cmuli(v, sign);
            case 6: return mk<C>(-r2 * (v.x + sg * v.y), r2 * (sg * v.x - v.y));
            case 9: return mk<C>(-c1 * v.x + sg * s1 * v.y, -c1 * v.y - sg * s1 * v.x);
            default: return v;
            }
        };
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            C t[4];
#pragma unroll
            for (int n2 = 0; n2 < 4; ++n2) t[n2] = tw(a[n2][k1], n2 * k1);
            Dft<4, C>::run(t, sign);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) x[k1 + 4 * k2] = t[k2];
        }
    }
};

// compiler-level ordering of LDS accesses between the lanes of one wave
__device__ __forceinline__ void wave_order()
{
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}


// Courant maxima of one level (src/courant.f90:209-275, magnetic branch with
// the Alfven damping terms valri2/valhi2): over this thread's points the
// maxima of or4 (cf^2 orho2 vr^2 + af^2 valr^2 / (valr + valri2)) and or2
// (cf^2 orho2 (vt^2 + vp^2)/sin^2 + af^2 valh^2 / (valh + valhi2)), valr =
// LFfac orho1 br^2, valh = LFfac orho1 (bt^2 + bp^2)/sin^2, orho2 = orho1^2;
// reduced over the wave, one atomic maximum per wave and quantity
template <typename T, int NP>
__device__ inline void courant_max(const CourArgs &c, int lev, int nlev, T o2, T or2, T or4, T orho1,
                                   T LFfac, const T (*u)[NP], const T (*b)[NP])
{
    const T cf2 = T(c.par[0]), af2 = T(c.par[1]);
    const T vri2 = T(c.par[2 + lev]), vhi2 = T(c.par[2 + nlev + lev]);
    const T orho2 = orho1 * orho1;
    T mr = T(0), mh = T(0);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const T vr = u[0][p], vt = u[1][p], vp = u[2][p];
        const T br = b[0][p], bt = b[1][p], bp = b[2][p];
        const T valr = br * br * LFfac * orho1;
        const T valr2 = valr * valr / (valr + vri2);
        mr = fmax(mr, or4 * (cf2 * (orho2 * vr * vr) + af2 * valr2));
        const T valh2 = (bt * bt + bp * bp) * LFfac * o2 * orho1;
        const T valh2m = valh2 * valh2 / (valh2 + vhi2);
        mh = fmax(mh, or2 * (cf2 * ((vt * vt + vp * vp) * o2 * orho2) + af2 * valh2m));
    }
    double dr = (double)mr, dh = (double)mh;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        dr = fmax(dr, __shfl_xor(dr, off, 64));
        dh = fmax(dh, __shfl_xor(dh, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(c.max + 2 * lev, (unsigned long long)__double_as_longlong(dr));
        atomicMax(c.max + 2 * lev + 1, (unsigned long long)__double_as_longlong(dh));
    }
}

// ---------------------------------------------------------------------------
// Generic LDS row FFTs (any length M = 2^a 3^b 4^c 5^d), used by the per-call
// API kernels and the batched transform jobs.
struct FFTArgs {
    int M, N, nfac, fac[16];
    const void *tw;   // exp(-2 pi i k / M), k < M (precision of the kernel)
    const void *twr;  // exp(-2 pi i k / N), k <= M
};

// one Stockham pass from `in` to `out` over G rows of M complex values
template <int R, typename C>
__device__ void pass_pp(const C *in, C *out, int G, int M, int L, const C *tw, int sign)
{
    const int T = M / R;
    const int tstride = M / (L * R);
    const int nthr = blockDim.x;
    for (int idx = threadIdx.x; idx < G * T; idx += nthr) {
        int g = idx / T, j = idx - g * T;
        int k = j % L;
        const C *ip = in + g * M;
        C x[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            C v = ip[j + s * T];
            if (s > 0) {
                C w = tw[s * k * tstride];
                if (sign > 0) w = cconj(w);
                v = cmul(v, w);
            }
            x[s] = v;
        }
        Dft<R, C>::run(x, sign);
        C *op = out + g * M + (j - k) * R + k;
#pragma unroll
        for (int s = 0; s < R; ++s) op[s * L] = x[s];
    }
}

// complex FFT of G rows held in buffer a (b is scratch); returns the buffer
// that holds the result
template <typename C>
__device__ C *fft_pp(C *a, C *b, int G, const FFTArgs &f, const C *tw, int sign)
{
    int L = 1;
    for (int s = 0; s < f.nfac; ++s) {
        int R = f.fac[s];
        switch (R) {
        case 4: pass_pp<4>(a, b, G, f.M, L, tw, sign); break;
        case 2: pass_pp<2>(a, b, G, f.M, L, tw, sign); break;
        case 3: pass_pp<3>(a, b, G, f.M, L, tw, sign); break;
        default: pass_pp<5>(a, b, G, f.M, L, tw, sign); break;
        }
        __syncthreads();
        C *t = a;
        a = b;
        b = t;
        L *= R;
    }
    return a;
}

// Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N} (X_k - conj X_{M-k}), real
// parts only for X_0 and X_M; X_k is read from global memory at X[k xs]
// (zero for k >= n_m), Z written to LDS row z
template <typename C>
__device__ void irfft_pre(const C *__restrict__ X, size_t xs, int n_m, C *z, int M, const C *twr,
                          int tid0, int nthr)
{
    typedef decltype(C::x) T;
    for (int k = tid0; k <= M / 2; k += nthr) {
        int k2 = M - k;
        C a = (k < n_m) ? X[(size_t)k * xs] : mk<C>(T(0), T(0));
        C c = (k2 < n_m) ? X[(size_t)k2 * xs] : mk<C>(T(0), T(0));
        if (k == 0) {
            a.y = T(0);
            c.y = T(0);
        }
        C za = cadd(cadd(a, cconj(c)), cmuli(cmul(cconj(twr[k]), csub(a, cconj(c))), 1));
        z[k] = za;
        if (k2 < M && k2 != k) {
            C zb = cadd(cadd(c, cconj(a)), cmuli(cmul(cconj(twr[k2]), csub(c, cconj(a))), 1));
            z[k2] = zb;
        }
    }
}

// X_k = (E_k + e^{-2 pi i k/N} O_k) * scale for k < n_m from the FFT of
// z_n = x_{2n} + i x_{2n+1}
template <typename C>
__device__ inline C rfft_post(const C *z, int k, int M, const C *twr, decltype(C::x) scale)
{
    typedef decltype(C::x) T;
    C zk = z[k == M ? 0 : k];
    C zc = cconj(z[k == 0 ? 0 : M - k]);
    C E = mk<C>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
    C D = csub(zk, zc);
    C O = mk<C>(T(0.5) * D.y, T(-0.5) * D.x);
    C X = cadd(E, cmul(twr[k], O));
    X.x *= scale;
    X.y *= scale;
    if (k == 0) X.y = T(0);
    return X;
}

// the same on a slot row (padded at 12 points per lane, see pidx)
template <int P, typename C>
__device__ inline C rfft_post_padded(const C *z, int k, int M, const C *twr, decltype(C::x) scale)
{
    typedef decltype(C::x) T;
    const int k1 = k == M ? 0 : k, k2 = k == 0 ? 0 : M - k;
    C zk = z[P == 12 ? k1 + (k1 >> 5) : k1];
    C zc = cconj(z[P == 12 ? k2 + (k2 >> 5) : k2]);
    C E = mk<C>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
    C D = csub(zk, zc);
    C O = mk<C>(T(0.5) * D.y, T(-0.5) * D.x);
    C X = cadd(E, cmul(twr[k], O));
    X.x *= scale;
    X.y *= scale;
    if (k == 0) X.y = T(0);
    return X;
}

template <typename C> __device__ void load_twiddles(C *s_tw, C *s_twr, const FFTArgs &f)
{
    const C *tw = reinterpret_cast<const C *>(f.tw);
    const C *twr = reinterpret_cast<const C *>(f.twr);
    for (int k = threadIdx.x; k < f.M; k += blockDim.x) s_tw[k] = tw[k];
    for (int k = threadIdx.x; k <= f.M; k += blockDim.x) s_twr[k] = twr[k];
}

// ---------------------------------------------------------------------------
// Wave-slot FFTs: a slot of Q lanes of one wave transforms one row of M
// points held in LDS, each lane holding P points.
// slots of 12 points per lane pad their rows by one complex per 32 (index
// i at i + i/32): the radix-12 pass writes 12 consecutive points per lane, a
// 12-point lane stride that otherwise maps every 8th lane to the same LDS
// banks (measured: 1.364 -> 1.327 ms at M = 384, 11.1 -> 10.5 ms at M = 768;
// at 24 points per lane the padding cost 4 %, so M = 1536 is not padded)
template <int P> __device__ __forceinline__ int pidx(int i) { return P == 12 ? i + (i >> 5) : i; }
template <int M, int P> constexpr int padded_row() { return P == 12 ? M + (M >> 5) + 1 : M; }

template <int R, int L, int M, int P, typename C>
__device__ __forceinline__ void slot_pass(C *buf, int t, const C *tw, int sign)
{
    constexpr int Q = M / P, NB = P / R;  // lanes per slot, butterflies per lane
    static_assert(P % R == 0 && Q * P == M, "pass does not tile the slot");
    constexpr int T = M / R, tstride = M / (L * R);
    // k = (t + Q b) mod L from t mod L (t < Q: no division when L >= Q)
    const int tm = L >= Q ? t : t % L;
    auto kof = [&](int b) {
        const int k = tm + (Q * b) % L;
        return k >= L ? k - L : k;
    };
    C x[NB][R];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + Q * b;
        const int k = kof(b);
#pragma unroll
        for (int s = 0; s < R; ++s) {
            C v = buf[pidx<P>(j + s * T)];
            if constexpr (L > 1) {
                if (s > 0) {
                    const int ti = s * k * tstride;
                    C w = tw[ti + (ti >> 3)];  // table padded by one entry per 8 (banks)
                    if (sign > 0) w = cconj(w);
                    v = cmul(v, w);
                }
            }
            x[b][s] = v;
        }
    }
    wave_order();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + Q * b;
        const int k = kof(b);
        Dft<R, C>::run(x[b], sign);
        const int o = (j - k) * R + k;
#pragma unroll
        for (int s = 0; s < R; ++s) buf[pidx<P>(o + s * L)] = x[b][s];
    }
    wave_order();
}

// M = 3 4^A 2^B = 12 4^(A-1) 2^B: a radix-12 pass (no twiddles), A - 1
// radix-4 passes, B radix-2 passes; Q = min(M / 12, 64) lanes of P points
constexpr int pow4(int a) { return a == 0 ? 1 : 4 * pow4(a - 1); }
template <int A, int B> struct SlotFFT {
    static constexpr int M = 3 * pow4(A) * (B ? 2 : 1);
    static constexpr int Q = M / 12 <= 64 ? M / 12 : 64, P = M / Q;
    // threads of the grid-stage block: six slots, at least one wave
    static constexpr int NTHR = (6 * Q < 64 ? 64 : 6 * Q + 63) / 64 * 64;
    static_assert(Q * P == M && P % 12 == 0 && 64 % Q == 0, "unsupported slot length");
    template <int I, typename C>
    __device__ __forceinline__ static void radix4(C *buf, int t, const C *tw, int sign)
    {
        if constexpr (I < A) {
            slot_pass<4, 12 * pow4(I - 1), M, P>(buf, t, tw, sign);
            radix4<I + 1>(buf, t, tw, sign);
        }
    }
    template <typename C> __device__ __forceinline__ static void run(C *buf, int t, const C *tw, int sign)
    {
        slot_pass<12, 1, M, P>(buf, t, tw, sign);
        radix4<1>(buf, t, tw, sign);
        if constexpr (B) slot_pass<2, 12 * pow4(A - 1), M, P>(buf, t, tw, sign);
    }
};

// the passes of a slot FFT of length M = 12 R1 R2 ... (Ri in 2, 3, 4): a
// radix-12 pass, then the listed radices in order, 12 points per lane
template <int M, int L, int R, int... Rest, typename C>
__device__ __forceinline__ void slot_chain(C *buf, int t, const C *tw, int sign)
{
    slot_pass<R, L, M, 12>(buf, t, tw, sign);
    if constexpr (sizeof...(Rest) > 0) slot_chain<M, L * R, Rest...>(buf, t, tw, sign);
}

}  // namespace

}  // namespace mhx
