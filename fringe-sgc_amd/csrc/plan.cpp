// Host-side motif planning: from a matching order, decide the core class and the
// reference kernel semantics that apply, and precompute every per-motif constant the
// kernels need (fringe vectors, rotations, degree thresholds, the functionX plan).
#include "fsgc_plan.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>

namespace fsgc {

namespace {
enum { iu = 1, iv = 2, iw = 4, iuv = 3, iuw = 5, ivw = 6, iuvw = 7 };  // FC:61-67

void put_row(int fr[6][8], int r, int uvw, int uv, int uw, int vw, int u, int v, int w)
{
  fr[r][0] = 0; fr[r][iuvw] = uvw; fr[r][iuv] = uv; fr[r][iuw] = uw; fr[r][ivw] = vw;
  fr[r][iu] = u; fr[r][iv] = v; fr[r][iw] = w;
}

// rotate: (uv, uw, vw, u, v, w) <- (vw, uv, uw, v, w, u) and its square
void rotations_from(int fr[6][8], int& p, int q)
{
  put_row(fr, p++, fr[q][iuvw], fr[q][ivw], fr[q][iuv], fr[q][iuw], fr[q][iv], fr[q][iw], fr[q][iu]);
  put_row(fr, p++, fr[q][iuvw], fr[q][iuw], fr[q][ivw], fr[q][iuv], fr[q][iw], fr[q][iu], fr[q][iv]);
}
}  // namespace

// FC:1291-1317: the distinct relabelings of a triangle's fringe vector (1, 3 or 6).
int triangle_rotations(const int f[8], int fr[6][8])
{
  int p = 0;
  put_row(fr, p++, f[iuvw], f[iuv], f[iuw], f[ivw], f[iu], f[iv], f[iw]);
  if ((f[iu] != f[iv]) || (f[iu] != f[iw]) || (f[iuv] != f[iuw]) || (f[iuv] != f[ivw])) {
    rotations_from(fr, p, 0);
    if (((f[iv] != f[iw]) || (f[iuv] != f[iuw])) && ((f[iu] != f[iw]) || (f[iuv] != f[ivw])) &&
        ((f[iu] != f[iv]) || (f[iuw] != f[ivw]))) {
      put_row(fr, p++, f[iuvw], f[iuw], f[iuv], f[ivw], f[iu], f[iw], f[iv]);
      rotations_from(fr, p, p - 1);
    }
  }
  return p;
}

// FC:1754-1780: a wedge always has the 3 rotations (used for triangles) and the flip
// v <-> w when the endpoints' fringe vectors differ.
int wedge_rotations(const int f[8], int fr[6][8])
{
  int p = 0;
  put_row(fr, p++, f[iuvw], f[iuv], f[iuw], f[ivw], f[iu], f[iv], f[iw]);
  rotations_from(fr, p, 0);
  if ((f[iv] != f[iw]) || (f[iuv] != f[iuw])) {
    put_row(fr, p++, f[iuvw], f[iuw], f[iuv], f[ivw], f[iu], f[iw], f[iv]);
    rotations_from(fr, p, p - 1);
  }
  return p;
}

int classify(const MatchingOrder& mo)
{
  if (mo.nodes == 1) return kVertex;
  if (mo.nodes == 2) return kEdge;
  if (mo.nodes == 3) return (__builtin_popcount(mo.list[1]) == 1) ? kWedge : kTriangle;  // FC:2250
  return kGeneral;
}

int validate_mo(const MatchingOrder& mo, std::string* err)
{
  if (mo.nodes < 1 || mo.nodes > kMaxCore) { if (err) *err = "matching order: core size must be 1..6"; return -22; }
  if (mo.anchorsets > 63) { if (err) *err = "matching order: too many anchor sets"; return -22; }
  const int full = (1 << mo.nodes) - 1;
  int total = mo.nodes;
  for (int i = 0; i < mo.anchorsets; i++) {
    if (mo.anch[i] == 0 || (mo.anch[i] & ~full)) { if (err) *err = "matching order: anchor set outside the core"; return -22; }
    if (i > 0 && mo.anch[i] <= mo.anch[i - 1]) { if (err) *err = "matching order: anchor sets not ascending"; return -22; }
    total += mo.fcnt[i];
  }
  if (total > kMaxPattern) { if (err) *err = "matching order: more than 32 motif vertices"; return -22; }
  for (int i = 0; i + 1 < mo.nodes; i++) {
    if (mo.list[i] == 0 || (mo.list[i] >> (i + 1))) { if (err) *err = "matching order: invalid connection list"; return -22; }
  }
  return 0;
}

// The reference's functionX repeats its first level fcnt[last] + 1 times when the last
// anchor set is the whole anchor universe and there is more than one anchor set
// (FC:428 computes `a` without the idx == top test of FC:480). Every general-core sum
// (data graph and motif graph) is scaled by that constant.
int general_quirk_factor(const MatchingOrder& mo)
{
  if (mo.nodes < 4 || mo.anchorsets < 2) return 1;
  int anchors = 0;
  for (int i = 0; i < mo.anchorsets; i++) anchors |= mo.anch[i];
  return (mo.anch[mo.anchorsets - 1] == anchors) ? mo.fcnt[mo.anchorsets - 1] + 1 : 1;
}

// Tables of the region DP for a general core: sets S_i = anch[i] with f_i fringes;
// the count of disjoint choices F_i within the union of the regions R >= S_i (the
// quantity functionX sums, FC:408-486) is, region by region,
//   new[t] = sum_{d <= t, d_i = 0 unless S_i <= R} old[t - d] * C(n_R, |d|) * |d|! / prod d_i!
// starting from old = [1, 0, ...] and read at t = (f_1, ..., f_A).
static void build_region_dp(const MatchingOrder& mo, int top, JobDesc* j)
{
  j->dp_ns = 0;
  j->dp_nreg = 0;
  const int A = mo.anchorsets;
  if (A < 1 || A > 8) return;
  long long ns = 1;
  for (int i = 0; i < A; i++) {
    if (mo.fcnt[i] < 1 || mo.fcnt[i] > 63) return;
    ns *= mo.fcnt[i] + 1;
    if (ns > 64) return;
  }
  for (int t = 0; t < ns; t++) {
    int rest = t, sum = 0;
    uint64_t dig = 0;
    unsigned __int128 mult = 1;
    for (int i = 0; i < A; i++) {
      const int ti = rest % (mo.fcnt[i] + 1);
      rest /= mo.fcnt[i] + 1;
      dig |= (uint64_t)ti << (8 * i);
      // multinomial as a product of binomials C(sum + ti, ti)
      for (int q = 1; q <= ti; q++) mult = mult * (unsigned)(sum + q) / (unsigned)q;
      sum += ti;
      if (mult >> 64) return;
    }
    j->dp_dig[t] = dig;
    j->dp_sum[t] = sum;
    j->dp_mult[t] = (uint64_t)mult;
  }
  int nr = 0;
  for (int R = 1; R <= top; R++) {
    if ((R & ~top) != 0) continue;
    uint64_t allow = 0;
    int maxj = 0;
    for (int i = 0; i < A; i++)
      if ((mo.anch[i] & ~R) == 0) { allow |= 0xffull << (8 * i); maxj += mo.fcnt[i]; }
    if (!allow) continue;
    j->dp_reg[nr] = (int8_t)R;
    j->dp_allow[nr] = allow;
    j->dp_maxj[nr] = std::min(maxj, 63);
    nr++;
  }
  j->dp_nreg = nr;
  j->dp_ns = (int)ns;
}

int make_job(const MatchingOrder& mo, JobDesc* j, std::string* err)
{
  int rc = validate_mo(mo, err);
  if (rc) return rc;
  memset(j, 0, sizeof(*j));
  j->cls = classify(mo);
  j->smod = 1;
  int f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  switch (j->cls) {
    case kVertex:  // FC:2172-2191
      if (mo.anchorsets == 0) j->kind = kVertexCount;
      else if (mo.fcnt[0] == 1) j->kind = kVertexEdges;
      else { j->kind = kVertexStars; j->p[0] = mo.fcnt[0]; }
      break;
    case kEdge: {  // FC:2106-2123
      for (int i = 0; i < mo.anchorsets; i++) f[mo.anch[i] & 3] = mo.fcnt[i];
      const int f_u = std::min(f[1], f[2]), f_uv = f[3], f_v = std::max(f[1], f[2]);
      const bool tri = (f_uv == 1) && ((f_u == 0) || (f_v == 0));
      j->kind = tri ? ((f_u == 0 && f_v == 0) ? kEdgeTriangle : kEdgeTailed) : kEdgeGeneral;
      j->p[0] = f_u; j->p[1] = f_uv; j->p[2] = f_v;
      j->p[3] = tri ? 1 : f_u + f_uv;
      j->p[4] = tri ? 1 : f_u + f_uv + f_v;
      break;
    }
    case kWedge: {  // FC:1783-1820
      for (int i = 0; i < mo.anchorsets; i++) f[mo.anch[i] & 7] = mo.fcnt[i];
      if (mo.list[1] == 2) { std::swap(f[iu], f[iv]); std::swap(f[iuw], f[ivw]); }
      const int min_u = f[iuvw] + f[iuv] + f[iuw] + f[iu] + 2;
      const int min_v = f[iuvw] + f[iuv] + f[ivw] + f[iv] + 1;
      const int min_w = f[iuvw] + f[iuw] + f[ivw] + f[iw] + 1;
      const int sum3 = min_u + min_v + min_w;
      const int sum2 = sum3 - std::max(std::max(min_u, min_v), min_w);
      const int sum1 = std::min(std::min(min_u, min_v), min_w);
      const int lim = std::min(min_v, min_w);
      j->p[0] = min_u; j->p[1] = lim; j->p[2] = sum1; j->p[3] = sum2; j->p[4] = sum3;
      if ((f[iuvw] == 0) && (f[iuv] == 0) && (f[iuw] == 0) && (f[iu] == 0)) {
        j->kind = kWedgeSpecial;  // FC:1409-1415
        j->p[5] = f[iv]; j->p[6] = f[ivw]; j->p[7] = f[iw];
        j->p[8] = f[iv] + f[ivw] + 1;
        j->p[9] = f[ivw] + f[iw] + 1;
        j->p[10] = std::min(j->p[8], j->p[9]);
      } else {
        j->kind = kWedgeGeneral;
        j->rot = wedge_rotations(f, j->fr);
        if (f[iuvw] == 0 && f[iuv] == 0 && f[iuw] == 0 && f[ivw] == 0 && f[iv] == 1 && f[iw] == 1) {
          j->p[11] = 1;
          j->p[12] = f[iu];
        }
      }
      break;
    }
    case kTriangle: {  // FC:1320-1368
      for (int i = 0; i < mo.anchorsets; i++) f[mo.anch[i] & 7] = mo.fcnt[i];
      const int min_u = f[iuvw] + f[iuv] + f[iuw] + f[iu];
      const int min_v = f[iuvw] + f[iuv] + f[ivw] + f[iv];
      const int min_w = f[iuvw] + f[iuw] + f[ivw] + f[iw];
      const int sum3 = min_u + min_v + min_w + 6;
      const int sum2 = sum3 - (std::max(std::max(min_u, min_v), min_w) + 2);
      const int sum1 = std::min(std::min(min_u, min_v), min_w) + 2;
      j->p[0] = sum1; j->p[1] = sum2; j->p[2] = sum3;
      j->p[3] = f[iu] + f[iv] + f[iw];
      j->rot = triangle_rotations(f, j->fr);
      if ((f[iuvw] == 1) && (f[iuv] + f[iuw] + f[ivw] == 0) && ((f[iu] == 0) + (f[iv] == 0) + (f[iw] == 0) >= 2))
        j->kind = (j->p[3] == 0) ? kTriClique : kTriTailedClique;
      else
        j->kind = kTriGeneral;
      break;
    }
    default: {  // FC:666-692
      int anchors = 0;
      for (int i = 0; i < mo.anchorsets; i++) anchors |= mo.anch[i];
      const int nonanch = anchors ^ ((1 << mo.nodes) - 1);
      const int top = ((1 << mo.nodes) - 1) ^ nonanch;
      j->p[0] = mo.nodes;
      j->p[1] = mo.startdeg;
      for (int i = 0; i < 5; i++) { j->p[2 + i] = mo.destdeg[i]; j->p[7 + i] = mo.list[i]; }
      j->p[12] = nonanch;
      j->p[13] = top;
      for (int i = 0; i < mo.anchorsets; i++) j->fcnt_by_mask[mo.anch[i]] = mo.fcnt[i];
      // evaluation plan: anchor sets last to first, each over its superset regions
      int L = 0;
      for (int asi = mo.anchorsets - 1; asi >= 0; asi--) {
        const int S = mo.anch[asi];
        int R = S;
        for (;;) {
          if (L >= kMaxPlanLevels) { if (err) *err = "motif too complex: fringe plan exceeds 96 levels"; return -7; }
          j->lev_set[L] = (int8_t)S;
          j->lev_reg[L] = (int8_t)R;
          j->lev_last[L] = (R == top);
          j->lev_final[L] = (R == top) && (asi == 0);
          L++;
          if (R == top) break;
          R = (((R | nonanch) + 1) | S) & ~nonanch;
        }
      }
      j->nlev = L;
      build_region_dp(mo, top, j);
      break;
    }
  }
  return 0;
}

// ---------------------------------------------------------------- inclusion-exclusion

namespace {

// The fringes as labelled items, item q anchored on the core-vertex mask lab[q]. By
// Moebius inversion over the set partitions P of the items, the placements of the
// items on distinct non-core vertices, each adjacent to all of its anchor set, number
//   sum_P prod_{B in P} (-1)^{|B|-1} (|B|-1)! * N[union of the anchor sets in B];
// returned as (sorted unions, coefficient) with equal multisets merged.
std::vector<std::pair<std::vector<int>, long long>> ie_terms(const std::vector<int>& lab)
{
  const int nf = (int)lab.size();
  std::map<std::vector<int>, long long> T;
  std::vector<int> blk(nf, 0);
  std::function<void(int, int)> rec = [&](int i, int nb) {
    if (i == nf) {
      std::vector<int> um(nb, 0), sz(nb, 0);
      for (int q = 0; q < nf; q++) { um[blk[q]] |= lab[q]; sz[blk[q]]++; }
      long long mu = 1;
      for (int b = 0; b < nb; b++) {
        long long fct = 1;
        for (int k = 2; k < sz[b]; k++) fct *= k;
        mu *= (sz[b] % 2 == 1) ? fct : -fct;
      }
      std::sort(um.begin(), um.end());
      T[um] += mu;
      return;
    }
    for (int b = 0; b <= nb; b++) {
      blk[i] = b;
      rec(i + 1, b == nb ? nb + 1 : nb);
    }
  };
  rec(0, 0);
  std::vector<std::pair<std::vector<int>, long long>> out;
  for (auto& kv : T)
    if (kv.second) out.push_back(kv);
  return out;
}

uint32_t ie_pack(const std::vector<int>& masks)
{
  uint32_t pk = (uint32_t)masks.size() << 28;
  for (size_t f = 0; f < masks.size(); f++) pk |= (uint32_t)masks[f] << (4 * f);
  return pk;
}

// den = 2^shift * odd; the inverse of odd modulo 2^128 by Newton's iteration
void ie_den(u128 den, int32_t* shift, uint64_t* lo, uint64_t* hi)
{
  int s = 0;
  while (((den >> s) & 1) == 0) s++;
  const u128 odd = den >> s;
  u128 inv = odd;
  for (int it = 0; it < 8; it++) inv = inv * ((u128)2 - odd * inv);
  *shift = s;
  *lo = (uint64_t)inv;
  *hi = (uint64_t)(inv >> 64);
}

// the reference's binomial C(n, k) (FC:189-197) wraps its u64 once k * C(n, k) >= 2^64:
// the smallest such n
int32_t nck_wrap_threshold(int k)
{
  auto wraps = [&](long long n) {
    u128 r = 1;
    for (int i = 1; i <= k; i++) {
      r = r * (u128)(n - k + i) / (u128)i;
      if (r >> 64) return true;
    }
    return ((r * (u128)k) >> 64) != 0;
  };
  long long lo = k, hi = (1ll << 31);
  while (lo < hi) {
    const long long mid = (lo + hi) / 2;
    if (wraps(mid)) hi = mid; else lo = mid + 1;
  }
  return (int32_t)std::min<long long>(lo, INT32_MAX);
}

bool maxdeg_pow_fits(int32_t maxdeg, int f)
{
  if (maxdeg < 0 || maxdeg >= (1 << 21)) return false;
  u128 p = 1;
  for (int i = 0; i < f; i++) {
    if (maxdeg > 0 && p > (~(u128)0) / (u128)(unsigned)maxdeg) return false;
    p *= (u128)(unsigned)maxdeg;
  }
  return true;
}

}  // namespace

// ---------------------------------------------------------------- 4-clique cores

bool clique_core(const JobDesc& j)
{
  return j.cls == kGeneral && j.p[0] == 4 && j.p[7] == 1 && j.p[8] == 3 && j.p[9] == 7;
}

int job_fringes(const JobDesc& j)
{
  int s = 0;
  for (int m = 1; m < 64; m++) s += j.fcnt_by_mask[m];
  return s;
}

// The clique schedule evaluates F * den modulo 2^128 and divides exactly, which is
// exact when den * F < 2^128. den * F <= prod_S N_S^{f_S} <= maxdeg^fringes; three
// factors of N < 2^21 are multiplied in 64 bits.
bool clique_fits(const JobDesc& j, int32_t maxdeg)
{
  if (!clique_core(j)) return false;
  const int f = job_fringes(j);
  if (f > kCliqueMaxFringes || !maxdeg_pow_fits(maxdeg, f)) return false;
  CliqueTab tab;
  std::vector<CliqueTerm> terms;
  return build_clique_tab(j, &tab, &terms, nullptr) == 0;
}

// Tables of a 4-clique core. The fringes are labelled items; fringe q must go to a
// non-core vertex adjacent to every core vertex of its anchor set S_q, all distinct.
// By Moebius inversion over the set partitions P of the items, the number of such
// injective placements is
//   sum_P prod_{B in P} (-1)^{|B|-1} (|B|-1)! * N[union of the S_q, q in B],
// and the count of fringe sets is that divided by prod fcnt! (the items of one set are
// interchangeable). This is the quantity functionX enumerates (FC:408-486) for a core
// whose Venn regions are all non-negative, as every clique's are. Terms with the
// same multiset of unions are merged. Each of the 24 assignments of the clique's
// vertices to the core positions maps the unions to clique-vertex masks; assignments
// giving the same mapped term list form one class.
int build_clique_tab(const JobDesc& j, CliqueTab* tab, std::vector<CliqueTerm>* terms, std::string* err)
{
  if (!clique_core(j)) { if (err) *err = "not a 4-clique core"; return -22; }
  std::vector<int> lab;  // anchor mask (core positions) of each labelled fringe
  std::vector<std::pair<int, int>> sets;  // (mask, fringes)
  u128 den = 1;
  for (int m = 1; m < 16; m++) {
    const int f = j.fcnt_by_mask[m];
    if (f <= 0) continue;
    sets.push_back({m, f});
    for (int q = 0; q < f; q++) { lab.push_back(m); den *= (u128)(q + 1); }
  }
  for (int m = 16; m < 64; m++)
    if (j.fcnt_by_mask[m]) { if (err) *err = "anchor set outside a 4-vertex core"; return -22; }
  const int nf = (int)lab.size();
  if (nf > kCliqueMaxFringes) { if (err) *err = "too many fringes for the clique schedule"; return -7; }
  const std::vector<std::pair<std::vector<int>, long long>> base = ie_terms(lab);
  std::memset(tab, 0, sizeof(*tab));
  tab->thr[0] = j.p[1];
  for (int i = 0; i < 3; i++) tab->thr[1 + i] = j.p[2 + i];
  tab->max_thr = std::max(std::max(tab->thr[0], tab->thr[1]), std::max(tab->thr[2], tab->thr[3]));
  int perm[4] = {0, 1, 2, 3};
  std::map<std::vector<std::pair<std::vector<int>, long long>>, int> cls_of;
  std::vector<std::vector<std::pair<std::vector<int>, long long>>> cls_terms;
  auto mapm = [&](int m) {
    int r = 0;
    for (int i = 0; i < 4; i++)
      if (m >> i & 1) r |= 1 << perm[i];
    return r;
  };
  int pi = 0;
  do {
    std::vector<std::pair<std::vector<int>, long long>> mt;
    for (auto& t : base) {
      std::vector<int> ms;
      for (int m : t.first) ms.push_back(mapm(m));
      std::sort(ms.begin(), ms.end());
      mt.push_back({ms, t.second});
    }
    std::sort(mt.begin(), mt.end());
    auto it = cls_of.find(mt);
    int c;
    if (it == cls_of.end()) {
      c = (int)cls_terms.size();
      cls_of[mt] = c;
      cls_terms.push_back(mt);
      int k = 0;
      for (auto& s : sets) {
        if (s.second < 3) continue;
        if (k < 8) tab->nck_mask[c][k] = (uint8_t)mapm(s.first);
        k++;
      }
    } else {
      c = it->second;
    }
    tab->cls_perms[c] |= 1u << pi;
    tab->root_perms[perm[0]] |= 1u << pi;
    for (int i = 0; i < 4; i++) tab->perm[pi][i] = (uint8_t)perm[i];
    pi++;
  } while (std::next_permutation(perm, perm + 4));
  tab->ncls = (int)cls_terms.size();
  // the sum over all 24 assignments, merged: what a clique whose every assignment passes
  // the degree filters adds
  {
    std::map<std::vector<int>, long long> all;
    for (int c = 0; c < tab->ncls; c++) {
      const long long mult = __builtin_popcount(tab->cls_perms[c]);
      for (auto& t : cls_terms[c]) all[t.first] += mult * t.second;
    }
    std::vector<std::pair<std::vector<int>, long long>> v;
    for (auto& kv : all)
      if (kv.second) v.push_back(kv);
    cls_terms.push_back(v);
  }
  terms->clear();
  for (int c = 0; c <= tab->ncls; c++) {
    tab->term_off[c] = (int)terms->size();
    for (auto& t : cls_terms[c]) {
      if (t.second > INT32_MAX || t.second < INT32_MIN) { if (err) *err = "clique term coefficient out of range"; return -7; }
      const uint32_t pk = ie_pack(t.first);
      terms->push_back(CliqueTerm{(int32_t)t.second, pk});
    }
  }
  tab->term_off[tab->ncls + 1] = (int)terms->size();
  tab->nterm = (int)terms->size();
  if (tab->nterm > kCliqueMaxTerms) { if (err) *err = "too many clique terms"; return -7; }
  ie_den(den, &tab->den_shift, &tab->den_inv_lo, &tab->den_inv_hi);
  int k = 0;
  for (auto& st : sets) {
    if (st.second < 3) continue;
    if (k < 8) tab->nck_thr[k] = nck_wrap_threshold(st.second);
    k++;
  }
  tab->nck_n = std::min(k, 8);
  return 0;
}

// FC:121-186 (matchingOrder2ECLgraph): core vertices in matching order, then the
// fringes of each anchor set in order, sorted lists.
HostGraph motif_graph(const MatchingOrder& mo)
{
  int n = mo.nodes;
  for (int i = 0; i < mo.anchorsets; i++) n += mo.fcnt[i];
  std::vector<std::vector<int>> nb(n);
  for (int i = 0; i < mo.nodes - 1; i++)
    for (int k = 0; k <= i; k++)
      if (mo.list[i] & (1 << k)) { nb[i + 1].push_back(k); nb[k].push_back(i + 1); }
  int v = mo.nodes;
  for (int i = 0; i < mo.anchorsets; i++)
    for (int k = 0; k < mo.fcnt[i]; k++, v++)
      for (int c = 0; c < kMaxCore; c++)
        if (mo.anch[i] & (1 << c)) { nb[c].push_back(v); nb[v].push_back(c); }
  HostGraph g;
  g.n = n;
  g.nindex.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    std::sort(nb[i].begin(), nb[i].end());
    g.nlist.insert(g.nlist.end(), nb[i].begin(), nb[i].end());
    g.nindex[i + 1] = (int32_t)g.nlist.size();
  }
  g.m = (int64_t)g.nlist.size();
  return g;
}

}  // namespace fsgc
