// Device-side exact arithmetic, graph helpers and fringe formulas shared by the HIP
// translation units (internal linkage in each).
#pragma once
#include <hip/hip_runtime.h>

#include "fsgc_job.h"

namespace fsgc {
namespace dev {

typedef unsigned __int128 u128;
typedef unsigned long long ull;

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kVennSlots = 16;  // general-core Venn vectors staged per wave

enum { iu = 1, iv = 2, iw = 4, iuv = 3, iuw = 5, ivw = 6, iuvw = 7 };

// inverses of the odd numbers 1, 3, ..., 63 modulo 2^128 (for exact division), by
// Newton iteration x <- x (2 - o x) at compile time
constexpr unsigned __int128 inv_odd(unsigned o)
{
  unsigned __int128 x = o;
  for (int it = 0; it < 8; it++) x = x * ((unsigned __int128)2 - (unsigned __int128)o * x);
  return x;
}
#define FSGC_INV_LO(k) (ull)(inv_odd(2 * (k) + 1))
#define FSGC_INV_HI(k) (ull)(inv_odd(2 * (k) + 1) >> 64)
#define FSGC_INV_ROW(M) M(0), M(1), M(2), M(3), M(4), M(5), M(6), M(7), M(8), M(9), M(10), M(11), M(12), M(13), \
  M(14), M(15), M(16), M(17), M(18), M(19), M(20), M(21), M(22), M(23), M(24), M(25), M(26), M(27), M(28), M(29), \
  M(30), M(31)
static __constant__ ull c_inv_lo[32] = {FSGC_INV_ROW(FSGC_INV_LO)};
static __constant__ ull c_inv_hi[32] = {FSGC_INV_ROW(FSGC_INV_HI)};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T bcast(T x, int src) { return __shfl(x, src, kWave); }

__device__ __forceinline__ int wave_sum(int x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// ------------------------------------------------------------------ exact arithmetic

__device__ __forceinline__ u128 mul_lo128(u128 a, u128 b)
{
  const ull al = (ull)a, ah = (ull)(a >> 64), bl = (ull)b, bh = (ull)(b >> 64);
  u128 p = (u128)al * bl;
  const ull cross = ah * bl + al * bh;
  return p + ((u128)cross << 64);
}

// a * b with overflow detection (flag kOvf128)
__device__ __forceinline__ u128 mulx(u128 a, u128 b, uint32_t& fl)
{
  const ull al = (ull)a, ah = (ull)(a >> 64), bl = (ull)b, bh = (ull)(b >> 64);
  const u128 p = (u128)al * bl;
  if ((ah | bh) == 0) return p;
  if (ah && bh) fl |= kOvf128;
  const u128 c = (u128)ah * bl + (u128)al * bh;
  if ((ull)(c >> 64)) fl |= kOvf128;
  const u128 r = p + ((u128)(ull)c << 64);
  if (r < p) fl |= kOvf128;
  return r;
}

__device__ __forceinline__ u128 addx(u128 a, u128 b, uint32_t& fl)
{
  const u128 r = a + b;
  if (r < a) fl |= kOvf128;
  return r;
}

// x / i for 1 <= i <= 63 when the division is exact: shift out the power of two, then
// multiply by the inverse of the odd part modulo 2^128.
__device__ __forceinline__ u128 div_exact(u128 x, int i)
{
  const int s = __builtin_ctz(i);
  const int o = i >> s;
  x >>= s;
  if (o == 1) return x;
  const u128 inv = ((u128)c_inv_hi[o >> 1] << 64) | c_inv_lo[o >> 1];
  return mul_lo128(x, inv);
}

// FC:189-197 semantics: k = min(k, n - k); k < 0 -> 0; k == 0 -> 1; otherwise the exact
// binomial. The reference's u64 loop wraps iff k * C(n, k) >= 2^64 (kRefNckInexact).
// Exact u64 division by a small i when it divides x: shift, then multiply by the
// inverse of the odd part modulo 2^64 (no integer divide instruction).
__device__ __forceinline__ ull div_exact64(ull x, int i)
{
  const int s = __builtin_ctz(i);
  const int o = i >> s;
  x >>= s;
  return (o == 1) ? x : x * c_inv_lo[o >> 1];
}

static __device__ __noinline__ u128 ncr_general(int n, int k, uint32_t& fl);

// Binomials for k <= 4 straight from the falling factorial (exact in u128 for n < 2^31)
// times the inverse of k! modulo 2^128; larger k iterate. Division-free.
static __device__ __noinline__ u128 ncr(int n, int k, uint32_t& fl)
{
  k = min(k, n - k);
  if (k < 0) return 0;
  if (k == 0) return 1;
  const ull N = (unsigned)n;
  if (k == 1) return N;
  if (k == 2) return (N * (N - 1)) >> 1;  // n(n-1) < 2^62; 2 * C(n, 2) never wraps u64
  if (k == 3) {
    const u128 p = (u128)(N * (N - 1)) * (N - 2);  // 3 * C(n, 3) = p / 2
    if ((p >> 65) != 0) fl |= kRefNckInexact;
    return mul_lo128(p >> 1, ((u128)c_inv_hi[1] << 64) | c_inv_lo[1]);
  }
  if (k == 4) {
    const u128 p = (u128)(N * (N - 1)) * (u128)((N - 2) * (N - 3));  // 4 * C(n, 4) = p / 6
    const u128 c = mul_lo128(p >> 3, ((u128)c_inv_hi[1] << 64) | c_inv_lo[1]);
    if ((c >> 62) != 0) fl |= kRefNckInexact;
    return c;
  }
  return ncr_general(n, k, fl);
}

static __device__ __noinline__ u128 ncr_general(int n, int k, uint32_t& fl)
{
  // all intermediates i * C(n-k+i, i) <= k * C(n, k): stay in u64 while they fit
  ull r = 1;
  int i = 1;
  for (; i <= k; i++) {
    const ull f = (ull)(unsigned)(n - k + i);
    if (__umul64hi(r, f) != 0) break;
    r = div_exact64(r * f, i);
  }
  u128 R = r;
  if (i <= k) {
    fl |= kRefNckInexact;
    for (; i <= k; i++) {
      R = mulx(R, (u128)(unsigned)(n - k + i), fl);
      R = div_exact(R, i);
    }
    return R;
  }
  if (__umul64hi(r, (ull)k) != 0) fl |= kRefNckInexact;
  return R;
}

// ------------------------------------------------------------------ graph helpers

struct G {
  const int* __restrict__ idx;
  const int* __restrict__ adj;
  __device__ __forceinline__ int beg(int u) const { return idx[u]; }
  __device__ __forceinline__ int end(int u) const { return idx[u + 1]; }
  __device__ __forceinline__ int deg(int u) const { return idx[u + 1] - idx[u]; }
};

// membership of x in the sorted range adj[l, r)
__device__ __forceinline__ bool contains(const int* __restrict__ a, int l, int r, int x)
{
  int lo = l, hi = r;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < r && a[lo] == x;
}

// |N(a) & N(b)|, lane-strided over the shorter list (wave-cooperative, uniform args)
static __device__ int coop_common(const G& g, int ab, int ae, int bb, int be)
{
  if (ae - ab > be - bb) { int t = ab; ab = bb; bb = t; t = ae; ae = be; be = t; }
  int c = 0;
  for (int i = ab + lane_id(); i < ae; i += kWave) c += contains(g.adj, bb, be, g.adj[i]);
  return wave_sum(c);
}

// owner of CSR position j, searching vertices [lo, hi]
__device__ __forceinline__ int owner_of(const int* __restrict__ idx, int lo, int hi, long long j)
{
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (idx[mid] <= j) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------------ accumulation

struct LaneAcc {
  ull lo = 0, hi = 0, matches = 0;
  uint32_t fl = 0;
  __device__ __forceinline__ void add(u128 x)
  {
    const u128 s = (((u128)hi << 64) | lo);
    const u128 r = s + x;
    if (r < s) fl |= kOvf128;
    lo = (ull)r;
    hi = (ull)(r >> 64);
  }
};

static __device__ void flush(LaneAcc& a, Accum* out)
{
  ull lo = a.lo, hi = a.hi, mt = a.matches;
  uint32_t fl = a.fl;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const ull olo = __shfl_xor(lo, o, kWave), ohi = __shfl_xor(hi, o, kWave);
    const ull nlo = lo + olo;
    const ull c = nlo < lo;
    const u128 nh = (u128)hi + ohi + c;
    if ((ull)(nh >> 64)) fl |= kOvf128;
    lo = nlo;
    hi = (ull)nh;
    mt += __shfl_xor(mt, o, kWave);
    fl |= __shfl_xor(fl, o, kWave);
  }
  if (lane_id() == 0 && (lo | hi | mt | fl)) {
    const ull old = atomicAdd((ull*)&out->lo, lo);
    const ull carry = (old + lo) < old;
    const ull add_hi = hi + carry;
    if (add_hi < hi) fl |= kOvf128;
    if (add_hi) {
      const ull oh = atomicAdd((ull*)&out->hi, add_hi);
      if (oh + add_hi < oh) fl |= kOvf128;
    }
    if (mt) atomicAdd((ull*)&out->matches, mt);
    if (fl) atomicOr(&out->flags, fl);
  }
  a = LaneAcc();
}

// ------------------------------------------------------------------ fringe formulas

// binomial with the reference's conventions (the same values as ncr), inline for
// k <= 2, and for k = 3 (n < 2^21) and k = 4 (n < 2^16), where the falling factorial
// fits u64 and the reference's u64 loop does not wrap: the division by k! is exact,
// a shift followed by the inverse of 3 modulo 2^64 (3 * 0xAA..AB = 1 mod 2^64)
__device__ __forceinline__ u128 binom(int n, int k, uint32_t& fl)
{
  constexpr ull kInv3 = 0xAAAAAAAAAAAAAAABull;
  if (k == 0) return n >= 0 ? 1 : 0;
  if (n < k) return 0;
  if (k == 1) return (ull)n;
  if (k == 2) return ((ull)n * (ull)(n - 1)) >> 1;
  if (k == 3 && n < (1 << 21)) {
    const ull N = (ull)n;
    return ((N * (N - 1) * (N - 2)) >> 1) * kInv3;
  }
  if (k == 4 && n < (1 << 16)) {
    const ull N = (ull)n;
    return ((N * (N - 1) * (N - 2) * (N - 3)) >> 3) * kInv3;
  }
  return ncr(n, k, fl);
}

// The same count as fc3 when only the single-vertex anchor sets carry fringes, in a
// collapsed form: let X be the core vertex with the most fringes and Y, Z the others.
// For fixed disjoint F_Y, F_Z the fringes of X can be any f_X-subset of the
// X-eligible vertices they leave free, so
//   fc3 = sum_{k1 <= f_Y, k2 <= f_Z} C(n_X - k1 - k2, f_X) * S1(k1, k2) * S2(k1, k2)
// where k1 (k2) counts F_Y (F_Z) inside X's neighbourhood,
//   S1 = sum_i C(xyz, i) C(xy, k1 - i) C(xz + xyz - i, k2)        (X-side choices)
//   S2 = sum_j C(yz, j) C(y, f_Y - k1 - j) C(z + yz - j, f_Z - k2)  (outside X)
// with i fringes of Y in region xyz and j in region yz (Z avoids those).
// Exact for non-negative region sizes; pinned against fc3 by the parity tests.
__device__ __noinline__ u128 fc3_collapsed(int fx, int fy, int fz, int x, int y, int z, int xy, int xz, int yz,
                                           int xyz, uint32_t& fl)
{
  const int nx = x + xy + xz + xyz;
  u128 count = 0;
#pragma unroll 1
  for (int k1 = 0; k1 <= fy; k1++) {
#pragma unroll 1
    for (int k2 = 0; k2 <= fz; k2++) {
      const u128 cx = binom(nx - k1 - k2, fx, fl);
      if (cx == 0) continue;
      u128 s1 = 0;
#pragma unroll 1
      for (int i = 0; i <= k1; i++)
        s1 = addx(s1, mulx(mulx(binom(xyz, i, fl), binom(xy, k1 - i, fl), fl), binom(xz + xyz - i, k2, fl), fl), fl);
      if (s1 == 0) continue;
      u128 s2 = 0;
#pragma unroll 1
      for (int j = 0; j <= fy - k1; j++)
        s2 = addx(s2, mulx(mulx(binom(yz, j, fl), binom(y, fy - k1 - j, fl), fl), binom(z + yz - j, fz - k2, fl), fl), fl);
      count = addx(count, mulx(mulx(cx, s1, fl), s2, fl), fl);
    }
  }
  return count;
}

// fc3_collapsed when f_Y, f_Z <= 1, written out (regions non-negative). With
// b_c = C(n_X - c, f_X):
//   f_Y = f_Z = 0:  b_0
//   f_Y = 1:        b_0 (y + yz) + b_1 (xy + xyz)                  (f_Z = 1 mirrors it)
//   f_Y = f_Z = 1:  b_0 [y (z + yz) + yz (z + yz - 1)]
//                 + b_1 [(xy + xyz)(z + yz) + (xz + xyz)(y + yz)]
//                 + b_2 [xy (xz + xyz) + xyz (xz + xyz - 1)]
// (the (k1, k2) terms of the collapsed sum). The bracketed factors stay below 2^40.
__device__ __noinline__ u128 fc3_small(int fx, int fy, int fz, int x, int y, int z, int xy, int xz, int yz, int xyz,
                                          uint32_t& fl)
{
  const int nx = x + xy + xz + xyz;
  const u128 b0 = binom(nx, fx, fl);
  if ((fy | fz) == 0) return b0;
  const u128 b1 = binom(nx - 1, fx, fl);
  if (fz == 0) return addx(mulx(b0, (u128)(ull)(y + yz), fl), mulx(b1, (u128)(ull)(xy + xyz), fl), fl);
  if (fy == 0) return addx(mulx(b0, (u128)(ull)(z + yz), fl), mulx(b1, (u128)(ull)(xz + xyz), fl), fl);
  const u128 b2 = binom(nx - 2, fx, fl);
  const long long zyz = z + yz, xzxyz = xz + xyz;
  const ull t0 = (ull)((long long)y * zyz + (long long)yz * (zyz - 1));
  const ull t1 = (ull)((long long)(xy + xyz) * zyz + xzxyz * (long long)(y + yz));
  const ull t2 = (ull)((long long)xy * xzxyz + (long long)xyz * (xzxyz - 1));
  return addx(addx(mulx(b0, (u128)t0, fl), mulx(b1, (u128)t1, fl), fl), mulx(b2, (u128)t2, fl), fl);
}

// C(n, k) for k <= 2 with the binom conventions (0 when n < k), as a u64
__device__ __forceinline__ ull sbin(int n, int k)
{
  if (n < k) return 0;
  if (k == 0) return 1;
  if (k == 1) return (ull)n;
  return ((ull)n * (ull)(n - 1)) >> 1;
}

// fc3_collapsed with f_Y = FY and f_Z = FZ (each at most 2) fixed at compile time:
// the loops unroll and the small binomials become inline u64 expressions; only the
// C(n_X - c, f_X) stay general (one per c = 0 .. FY + FZ).
template <int FY, int FZ>
__device__ __noinline__ u128 fc3_fixed(int fx, int x, int y, int z, int xy, int xz, int yz, int xyz, uint32_t& fl)
{
  const int nx = x + xy + xz + xyz;
  u128 b[FY + FZ + 1];
#pragma unroll
  for (int c = 0; c <= FY + FZ; c++) b[c] = binom(nx - c, fx, fl);
  u128 count = 0;
#pragma unroll
  for (int k1 = 0; k1 <= FY; k1++) {
#pragma unroll
    for (int k2 = 0; k2 <= FZ; k2++) {
      u128 s1 = 0, s2 = 0;
#pragma unroll
      for (int i = 0; i <= k1; i++)
        s1 = addx(s1, mulx((u128)sbin(xyz, i) * sbin(xy, k1 - i), (u128)sbin(xz + xyz - i, k2), fl), fl);
#pragma unroll
      for (int j = 0; j <= FY - k1; j++)
        s2 = addx(s2, mulx((u128)sbin(yz, j) * sbin(y, FY - k1 - j), (u128)sbin(z + yz - j, FZ - k2), fl), fl);
      count = addx(count, mulx(mulx(b[k1 + k2], s1, fl), s2, fl), fl);
    }
  }
  return count;
}

// C(n, k) as a u64 with the binom conventions, for the cases binom evaluates inline:
// k <= 2, k = 3 with n < 2^21, k = 4 with n < 2^16 (the caller checks the range)
__device__ __forceinline__ ull ubin_small(int n, int k)
{
  constexpr ull kInv3 = 0xAAAAAAAAAAAAAAABull;
  if (k == 0) return n >= 0 ? 1 : 0;
  if (n < k) return 0;
  const ull N = (ull)n;
  if (k == 1) return N;
  if (k == 2) return (N * (N - 1)) >> 1;
  if (k == 3) return ((N * (N - 1) * (N - 2)) >> 1) * kInv3;
  return ((N * (N - 1) * (N - 2) * (N - 3)) >> 3) * kInv3;
}

// fc3_small inline, when its three binomials C(n_X - c, f_X) are inline u64 values
// (f_X <= 2, or f_X = 3 with n_X < 2^21, or f_X = 4 with n_X < 2^16). Then every
// binomial is below 2^61 and every bracket below 2^63, so the three products and
// their sum stay below 2^126: no flag can arise, and the value is fc3_small's. This
// keeps the common wedge and triangle cases free of a call, whose register saves and
// restores would go through scratch memory once per core match.
__device__ __forceinline__ bool fc3_small_inline_ok(int fx, int nx)
{
  return fx <= 2 || (fx == 3 && nx < (1 << 21)) || (fx == 4 && nx < (1 << 16));
}

__device__ __forceinline__ u128 fc3_small_inline(int fx, int fy, int fz, int nx, int y, int z, int xy, int xz, int yz,
                                                 int xyz)
{
  const ull b0 = ubin_small(nx, fx);
  if ((fy | fz) == 0) return b0;
  const ull b1 = ubin_small(nx - 1, fx);
  if (fz == 0) return (u128)b0 * (ull)(y + yz) + (u128)b1 * (ull)(xy + xyz);
  if (fy == 0) return (u128)b0 * (ull)(z + yz) + (u128)b1 * (ull)(xz + xyz);
  const ull b2 = ubin_small(nx - 2, fx);
  const long long zyz = z + yz, xzxyz = xz + xyz;
  const ull t0 = (ull)((long long)y * zyz + (long long)yz * (zyz - 1));
  const ull t1 = (ull)((long long)(xy + xyz) * zyz + xzxyz * (long long)(y + yz));
  const ull t2 = (ull)((long long)xy * xzxyz + (long long)xyz * (xzxyz - 1));
  return (u128)b0 * t0 + (u128)b1 * t1 + (u128)b2 * t2;
}

__device__ __forceinline__ u128 fc3_xyz(int fx, int fy, int fz, int x, int y, int z, int xy, int xz, int yz, int xyz,
                                        uint32_t& fl)
{
  if ((x | y | z | xy | xz | yz | xyz) >= 0) {
    if (fy <= 1 && fz <= 1) {
      const int nx = x + xy + xz + xyz;
      if (fc3_small_inline_ok(fx, nx)) return fc3_small_inline(fx, fy, fz, nx, y, z, xy, xz, yz, xyz);
      return fc3_small(fx, fy, fz, x, y, z, xy, xz, yz, xyz, fl);
    }
    if (fy <= 2 && fz <= 2) {
      switch (fy * 3 + fz) {
        case 2: return fc3_fixed<0, 2>(fx, x, y, z, xy, xz, yz, xyz, fl);
        case 5: return fc3_fixed<1, 2>(fx, x, y, z, xy, xz, yz, xyz, fl);
        case 6: return fc3_fixed<2, 0>(fx, x, y, z, xy, xz, yz, xyz, fl);
        case 7: return fc3_fixed<2, 1>(fx, x, y, z, xy, xz, yz, xyz, fl);
        default: return fc3_fixed<2, 2>(fx, x, y, z, xy, xz, yz, xyz, fl);
      }
    }
  }
  return fc3_collapsed(fx, fy, fz, x, y, z, xy, xz, yz, xyz, fl);
}

// A wedge core with f fringes on its centre X and one on each endpoint Y, Z (single-
// vertex anchor sets) has the fringe count fc3 = V - b0 (A + B) + (b0 - b2) s, where
//   V = b0 [(A - P)(B - Q) - R + A + B] + b1 [P (B - Q) + Q (A - P)] + b2 P Q,
// b_c = C(n_X - c, f), n_X = |N(X) minus the core|, A = |N(Y) minus the core|,
// B = |N(Z) minus the core|, P = |A & N(X)|, Q = |B & N(X)|, R = |A & B| and s the
// triple region |N(X) & N(Y) & N(Z)| (fc3_small's three terms regrouped by s). The
// count is linear in s with a coefficient of the centre alone, so a sum over wedges
// takes the s and A + B terms per centre and V per wedge. Every bracket is a count of
// pairs (R <= min(A, B)), so V >= 0; each stays below 2^63. With inline u64 binomials
// (fc3_small_inline_ok) V < 2^126 and no flag can arise.
__device__ __forceinline__ u128 wedge_lin_v(int f, int nx, int a, int b, int p, int q, int r, uint32_t& fl)
{
  const ull t0 = (ull)((long long)(a - p) * (b - q) - r + a + b);
  const ull t1 = (ull)((long long)p * (b - q) + (long long)q * (a - p));
  const ull t2 = (ull)((long long)p * q);
  if (fc3_small_inline_ok(f, nx))
    return (u128)ubin_small(nx, f) * t0 + (u128)ubin_small(nx - 1, f) * t1 + (u128)ubin_small(nx - 2, f) * t2;
  return addx(addx(mulx(binom(nx, f, fl), (u128)t0, fl), mulx(binom(nx - 1, f, fl), (u128)t1, fl), fl),
              mulx(binom(nx - 2, f, fl), (u128)t2, fl), fl);
}

__device__ __forceinline__ u128 fc3_singletons(const int* fr, const int* venn, uint32_t& fl)
{
  // X = the core vertex with the most fringes (ties: u, then v), Y and Z the others in
  // core order; the arguments are selected first so that fc3_xyz is expanded once
  const int fu = fr[iu], fv = fr[iv], fw = fr[iw];
  const int U = venn[iu], V = venn[iv], W = venn[iw], UV = venn[iuv], UW = venn[iuw], VW = venn[ivw];
  int fx, fy, fz, x, y, z, xy, xz, yz;
  if (fu >= fv && fu >= fw) {
    fx = fu; fy = fv; fz = fw; x = U; y = V; z = W; xy = UV; xz = UW; yz = VW;
  } else if (fv >= fw) {
    fx = fv; fy = fu; fz = fw; x = V; y = U; z = W; xy = UV; xz = VW; yz = UW;
  } else {
    fx = fw; fy = fu; fz = fv; x = W; y = U; z = V; xy = UW; xz = VW; yz = UV;
  }
  return fc3_xyz(fx, fy, fz, x, y, z, xy, xz, yz, venn[iuvw], fl);
}

__device__ __forceinline__ bool fc3_is_singleton(const int* fr) { return (fr[iuv] | fr[iuw] | fr[ivw] | fr[iuvw]) == 0; }

// fc3 when exactly one pair anchor set ps (iuv, iuw or ivw) carries fringes, besides
// single-vertex and all-three anchor sets: fc3_peeled's loops reduce to the placements
// of its f fringes, i in the triple region and f - i in the pair's own region, each
// times the single-vertex count on what remains. The same binomials in the same cases
// (so the same value and flags), inline. Regions non-negative.
__device__ __forceinline__ u128 fc3_one_pair(const int* fr, const int* venn, int ps, uint32_t& fl)
{
  int v[8];
#pragma unroll
  for (int r = 0; r < 8; r++) v[r] = venn[r];
  if (v[iuvw] < fr[iuvw]) return 0;
  const u128 cw = binom(v[iuvw], fr[iuvw], fl);
  const int s0 = v[iuvw] - fr[iuvw], p0 = v[ps], fp = fr[ps];
  u128 total = 0;
  for (int i = 0; i <= min(s0, fp); i++) {
    if (p0 < fp - i) continue;
    const u128 c = mulx(mulx(cw, binom(s0, i, fl), fl), binom(p0, fp - i, fl), fl);
    v[iuvw] = s0 - i;
    v[ps] = p0 - (fp - i);
    total = addx(total, mulx(c, fc3_singletons(fr, v, fl), fl), fl);
  }
  return total;
}

// Single-vertex anchor sets plus k = fr[iuvw] fringes anchored on all three core
// vertices: those can only take vertices of the triple region, so fc3 =
// C(s, k) * (the single-vertex count with the triple region reduced by k), which is
// fc3_peeled's one placement when no pair anchor set has fringes (the same binomials,
// so the same flags), inline. Regions non-negative.
__device__ __forceinline__ u128 fc3_singletons_uvw(const int* fr, const int* venn, uint32_t& fl)
{
  const int k = fr[iuvw];
  if (k == 0) return fc3_singletons(fr, venn, fl);
  if (venn[iuvw] < k) return 0;
  int v[8];
#pragma unroll
  for (int r = 0; r < 8; r++) v[r] = venn[r];
  const u128 c = binom(v[iuvw], k, fl);
  v[iuvw] -= k;
  return mulx(c, fc3_singletons(fr, v, fl), fl);
}

__device__ __forceinline__ bool venn_nonneg(const int* venn)
{
  return (venn[iu] | venn[iv] | venn[iw] | venn[iuv] | venn[iuw] | venn[ivw] | venn[iuvw]) >= 0;
}

// fc3 with multi-vertex anchor sets, non-negative regions: the reference's outer
// loops place the uvw, uv, uw and vw fringes (FC:249-276) and, for each placement,
// the single-vertex fringes are counted in closed or collapsed form on the regions
// that remain (instead of the four inner loops, FC:277-318).
static __device__ __noinline__ u128 fc3_peeled(const int* fr, const int* venn, uint32_t& fl)
{
  int v[8];
#pragma unroll
  for (int r = 0; r < 8; r++) v[r] = venn[r];
  if (v[iuvw] < fr[iuvw]) return 0;
  const u128 cuvw = binom(v[iuvw], fr[iuvw], fl);
  v[iuvw] -= fr[iuvw];
  u128 total = 0;
  for (int i = 0; i <= min(v[iuvw], fr[iuv]); i++) {
    if (v[iuv] < fr[iuv] - i) continue;
    const u128 c1 = mulx(mulx(cuvw, binom(v[iuvw], i, fl), fl), binom(v[iuv], fr[iuv] - i, fl), fl);
    v[iuvw] -= i; v[iuv] -= fr[iuv] - i;
    for (int j = 0; j <= min(v[iuvw], fr[iuw]); j++) {
      if (v[iuw] < fr[iuw] - j) continue;
      const u128 c2 = mulx(mulx(c1, binom(v[iuvw], j, fl), fl), binom(v[iuw], fr[iuw] - j, fl), fl);
      v[iuvw] -= j; v[iuw] -= fr[iuw] - j;
      for (int k = 0; k <= min(v[iuvw], fr[ivw]); k++) {
        if (v[ivw] < fr[ivw] - k) continue;
        const u128 c3 = mulx(mulx(c2, binom(v[iuvw], k, fl), fl), binom(v[ivw], fr[ivw] - k, fl), fl);
        v[iuvw] -= k; v[ivw] -= fr[ivw] - k;
        total = addx(total, mulx(c3, fc3_singletons(fr, v, fl), fl), fl);
        v[iuvw] += k; v[ivw] += fr[ivw] - k;
      }
      v[iuvw] += j; v[iuw] += fr[iuw] - j;
    }
    v[iuvw] += i; v[iuv] += fr[iuv] - i;
  }
  return total;
}

// FC:233-327 (fc3): ways to place the fringes of the 7 anchor sets of a 3-vertex core
// disjointly into the 7 Venn regions (loop order uvw, uv, uw, vw, u; v and w merged).
static __device__ __noinline__ u128 fc3(const int* fr, const int* venn, uint32_t& fl)
{
  const int fr_uvw = fr[iuvw], fr_uv = fr[iuv], fr_uw = fr[iuw], fr_vw = fr[ivw];
  const int fr_u = fr[iu], fr_v = fr[iv], fr_w = fr[iw];
  int vd_uvw = venn[iuvw], vd_uv = venn[iuv], vd_uw = venn[iuw], vd_vw = venn[ivw];
  const int vd_u = venn[iu], vd_v = venn[iv], vd_w = venn[iw];
  if (vd_uvw < fr_uvw) return 0;
  u128 count = 0;
  const u128 cuvw = ncr(vd_uvw, fr_uvw, fl);
  vd_uvw -= fr_uvw;
  for (int i = min(vd_uvw, fr_uv); i >= 0; i--) {
    if (vd_uv < fr_uv - i) break;
    const u128 c1 = mulx(mulx(cuvw, ncr(vd_uvw, i, fl), fl), ncr(vd_uv, fr_uv - i, fl), fl);
    vd_uvw -= i; vd_uv -= fr_uv - i;
    for (int j = min(vd_uvw, fr_uw); j >= 0; j--) {
      if (vd_uw < fr_uw - j) break;
      const u128 c2 = mulx(mulx(c1, ncr(vd_uvw, j, fl), fl), ncr(vd_uw, fr_uw - j, fl), fl);
      vd_uvw -= j; vd_uw -= fr_uw - j;
      for (int k = min(vd_uvw, fr_vw); k >= 0; k--) {
        if (vd_vw < fr_vw - k) break;
        const u128 c3 = mulx(mulx(c2, ncr(vd_uvw, k, fl), fl), ncr(vd_vw, fr_vw - k, fl), fl);
        vd_uvw -= k; vd_vw -= fr_vw - k;
      #pragma unroll 1
  for (int a = min(vd_uvw, fr_u); a >= 0; a--) {
          const u128 ta = mulx(c3, ncr(vd_uvw, a, fl), fl);
          vd_uvw -= a;
          int vd_uvw_vw = vd_uvw + vd_vw;
          for (int b = min(vd_uv, fr_u - a); b >= 0; b--) {
            const u128 tb = mulx(ta, ncr(vd_uv, b, fl), fl);
            vd_uv -= b;
            for (int c = min(vd_uw, fr_u - a - b); c >= 0; c--) {
              if (vd_u < fr_u - a - b - c) break;
              const u128 tc = mulx(mulx(tb, ncr(vd_uw, c, fl), fl), ncr(vd_u, fr_u - a - b - c, fl), fl);
              vd_uw -= c;
              for (int d = min(vd_uvw_vw, fr_v); d >= 0; d--) {
                if (vd_uv + vd_v < fr_v - d) break;
                const u128 td = mulx(mulx(tc, ncr(vd_uvw_vw, d, fl), fl), ncr(vd_uv + vd_v, fr_v - d, fl), fl);
                vd_uvw_vw -= d;
                if (vd_uvw_vw + vd_uw + vd_w >= fr_w)
                  count = addx(count, mulx(td, ncr(vd_uvw_vw + vd_uw + vd_w, fr_w, fl), fl), fl);
                vd_uvw_vw += d;
              }
              vd_uw += c;
            }
            vd_uv += b;
          }
          vd_uvw += a;
        }
        vd_uvw += k; vd_vw += fr_vw - k;
      }
      vd_uvw += j; vd_uw += fr_uw - j;
    }
    vd_uvw += i; vd_uv += fr_uv - i;
  }
  return count;
}

// FC:2030-2048 (general_edge_core) and FC:1381-1406 (wedge): the two-vertex closed
// form; `corr` is how many core vertices each list holds besides the shared ones
// (1 for an edge core, 1 or 2 for wedge endpoints).
static __device__ __noinline__ u128 pair_formula(int da, int db, int c_ab, int corr, int f_a, int f_ab, int f_b, uint32_t& fl)
{
  const int a_ab = c_ab - f_ab;
  if (a_ab < 0) return 0;
  const int ca = da - (corr + c_ab), cb = db - (corr + c_ab);
  const int bound = min(f_a, a_ab);
  u128 acc = 0;
  for (int j = 0; j <= bound; j++) {
    const int k = f_a - j;
    u128 term = mulx(binom(ca, k, fl), binom(a_ab + cb - j, f_b, fl), fl);
    if (f_a != f_b) term = addx(term, mulx(binom(cb, k, fl), binom(a_ab + ca - j, f_b, fl), fl), fl);
    acc = addx(acc, mulx(term, binom(a_ab, j, fl), fl), fl);
  }
  return mulx(acc, binom(c_ab, f_ab, fl), fl);
}

}  // namespace dev
}  // namespace fsgc
