// POD descriptors shared by host code and the HIP kernels.
#pragma once
#include <cstdint>

namespace fsgc {

enum MotifClass : int32_t { kVertex = 0, kEdge = 1, kWedge = 2, kTriangle = 3, kGeneral = 4 };

// Variants inside a class (which reference kernel's semantics apply).
enum : int32_t {
  kVertexCount = 0,    // anchorsets == 0 -> n           (FC:2175-2177)
  kVertexEdges = 1,    // fcnt[0] == 1 -> m / 2          (FC:2180-2182)
  kVertexStars = 2,    // sum_u C(d_u, t)                (FC:2152-2169)
  kEdgeTriangle = 0,   // triangles w < v < u            (FC:1864-1894)
  kEdgeTailed = 1,     // tailed triangles               (FC:1931-1966)
  kEdgeGeneral = 2,    // closed form per edge           (FC:1999-2051)
  kWedgeSpecial = 0,   // centre without fringes         (FC:1409-1477)
  kWedgeGeneral = 1,   // fc3 with rotations             (FC:1513-1650)
  kTriClique = 0,      // 4-cliques                      (FC:774-809)
  kTriTailedClique = 1,// tailed 4-cliques               (FC:891-931)
  kTriGeneral = 2,     // fc3 with rotations             (FC:1027-1062)
};

// Result flags (also in include/fsgc.h).
enum : uint32_t {
  kOvf64 = 1u << 0,
  kOvf128 = 1u << 1,
  kRefNckInexact = 1u << 2,
  kAutRemainder = 1u << 3,
  kRefAutTrunc = 1u << 4,
};

constexpr int kMaxPlanLevels = 96;

// Device-side u128 accumulator of one job (one motif on one graph).
struct Accum {
  uint64_t lo;
  uint64_t hi;
  uint64_t matches;
  uint32_t flags;
  uint32_t pad;
  uint64_t bytes;  // algorithmic global-memory bytes of the launch's schedule (pivot kernel)
};

// A term of an inclusion-exclusion fringe count (4-clique cores):
//   F * den = sum_t coef_t * prod_f N[mask_tf],
// where N[U] is the number of non-core vertices adjacent to every core vertex in U and
// den = prod over anchor sets of fcnt!. A term packs its factor count in bits 28..31
// and up to seven 4-bit core-vertex masks below.
struct CliqueTerm {
  int32_t coef;
  uint32_t packed;
};

struct JobDesc {
  int32_t cls;
  int32_t kind;
  const int32_t* idx;
  const int32_t* adj;
  int32_t n;
  int32_t m;
  int32_t smod;   // slice: only roots u with u % smod == srem (smod <= 1: all)
  int32_t srem;
  // scalar parameters; meaning per class:
  //   vertex:   p[0] = tails
  //   edge:     p[0] f_u, p[1] f_uv, p[2] f_v, p[3] lim1, p[4] lim2
  //   wedge:    p[0] min_u, p[1] lim, p[2] sum1, p[3] sum2, p[4] sum3,
  //             special: p[5] f_v, p[6] f_vw, p[7] f_w, p[8] smin_v, p[9] smin_w, p[10] slim
  //             general: p[11] = 1 when the count is linear in the triple region (f_u >= 1
  //             fringes on the centre, one on each endpoint, no multi-vertex anchor set;
  //             the pivot schedule then evaluates it without c3), p[12] f_u
  //   triangle: p[0] sum1, p[1] sum2, p[2] sum3, p[3] tails
  //   general:  p[0] nodes, p[1] startdeg, p[2..6] destdeg, p[7..11] list, p[12] nonanch, p[13] top
  int32_t p[16];
  int32_t rot;
  int32_t fr[6][8];
  // general core: fringe counts by anchor bitmask and the functionX evaluation plan
  int32_t fcnt_by_mask[64];
  int32_t nlev;                    // number of (anchor set, region) levels
  int8_t lev_set[kMaxPlanLevels];  // anchor bitmask of the level
  int8_t lev_reg[kMaxPlanLevels];  // region bitmask of the level
  int8_t lev_last[kMaxPlanLevels]; // 1 if the region is the last (top) one for this set
  int8_t lev_final[kMaxPlanLevels];// 1 if this is the last level of the whole plan
  // general core, the same count as a dynamic programme over the Venn regions
  // (dp_ns > 0): a state is the number of fringes already placed per anchor set,
  // numbered in mixed radix (set i has weight prod_{j<i} (f_j + 1))
  int32_t dp_ns;                   // states (<= 64); 0: evaluate by the plan above
  int32_t dp_nreg;                 // regions some anchor set may draw from
  int8_t dp_reg[64];               // their anchor bitmasks
  int32_t dp_maxj[64];             // most fringes a region can give (sum of f_i of the sets it serves)
  int32_t dp_sum[64];              // per state: fringes placed
  uint64_t dp_dig[64];             // per state: the count of set i in byte i
  uint64_t dp_mult[64];            // per state: (sum t_i)! / prod t_i!
  uint64_t dp_allow[64];           // per region: 0xff in byte i if set i may draw from it
  Accum* out;
};

constexpr int kCliqueMaxTerms = 8192;  // terms of all assignment classes of one motif
constexpr int kCliqueMaxFringes = 7;

// The 24 assignments of the clique's vertices (ordered by rank) to the motif's core
// positions, grouped into classes with equal fringe counts (assignments related by a
// symmetry of the anchored motif); per class the terms of its first assignment.
struct CliqueTab {
  int32_t thr[4];          // minimum degree per core position (startdeg, destdeg[0..2])
  int32_t max_thr;         // the largest of them: a clique of smaller degrees all above it passes every assignment
  int32_t ncls;
  int32_t nterm;
  int32_t use_all;         // the merged table is exact on this graph (24 * maxdeg^fringes < 2^128)
  uint32_t cls_perms[24];  // per class: bit i set for assignment i
  uint8_t perm[24][4];     // assignment i: core position -> clique vertex (0..3)
  uint32_t root_perms[4];  // per clique vertex: the assignments that put it at position 0 (the root)
  int32_t term_off[26];    // terms of class c: [term_off[c], term_off[c + 1]); class ncls is the
                           // sum over all 24 assignments (a clique that passes every filter)
  int32_t den_shift;       // den = 2^den_shift * odd
  uint64_t den_inv_lo;     // inverse of odd modulo 2^128
  uint64_t den_inv_hi;
  // anchor sets with at least three fringes: the value of N at which the reference's
  // u64 binomial of that many fringes wraps (kRefNckInexact), and per class the
  // set's clique-vertex mask
  int32_t nck_n;
  int32_t nck_thr[8];
  uint8_t nck_mask[24][8];
  const CliqueTerm* terms;  // device copy of the nterm terms
};

// One unit of the device work queue: a range of roots of one job.
// Roots are vertices for the vertex class and CSR positions otherwise.
struct WorkItem {
  int32_t job;
  int32_t u0;     // owner of position `begin` (edge-rooted classes)
  int32_t u1;     // owner of position `end - 1`
  // general cores: 0, or (parts << 8) | part for an item that holds one part of its
  // root edges' first-level candidates (those at list offset i with i % parts == part)
  int32_t split;
  int64_t begin;
  int64_t end;
};

}  // namespace fsgc
