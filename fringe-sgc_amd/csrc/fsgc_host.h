// Host-side types shared by the library, the CLIs and the device driver code.
#pragma once
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

namespace fsgc {

typedef unsigned __int128 u128;
typedef uint64_t u64;

constexpr int kMaxCore = 6;      // fringes.h:44 MaxCoreSize
constexpr int kMaxPattern = 32;  // patterns.h:45 MaxPatternSize

// The on-disk matching order: 139 bytes, no padding (fringes.h:50-58).
#pragma pack(push, 1)
struct MatchingOrder {
  uint8_t nodes;
  uint8_t anchorsets;
  uint8_t startdeg;
  uint8_t destdeg[kMaxCore - 1];
  uint8_t list[kMaxCore - 1];
  uint8_t anch[(1 << kMaxCore) - 1];
  uint8_t fcnt[(1 << kMaxCore) - 1];
};
#pragma pack(pop)
static_assert(sizeof(MatchingOrder) == 139, "matching order must be 139 bytes");

// Host CSR graph in the ECLgraph layout (ECLgraph.h:45-51), eweight dropped.
struct HostGraph {
  int32_t n = 0;
  int64_t m = 0;  // directed entries (ECLgraph.edges)
  std::vector<int32_t> nindex;
  std::vector<int32_t> nlist;
};

// Small undirected pattern graph (<= 32 vertices); adjacency as bit masks.
struct Pattern {
  std::string name;
  int nodes = 0;
  uint32_t adj[kMaxPattern] = {0};
  void add_edge(int a, int b) { if (a != b) { adj[a] |= 1u << b; adj[b] |= 1u << a; } }
  int degree(int v) const { return __builtin_popcount(adj[v]); }
};

// formats.cpp
int read_egr(const char* path, HostGraph* g, std::string* err);
int write_egr(const char* path, const HostGraph& g, std::string* err);
int validate_csr(const HostGraph& g, std::string* err);
void print_matching_order(const MatchingOrder& mo, std::string* out);

// preprocess.cpp
int predefined_pattern(int number, Pattern* p);
int pattern_from_edges(int nodes, int nedges, const int* edges, Pattern* p, std::string* err);
int preprocess(const Pattern& p, MatchingOrder* mo, std::string* log, std::string* dot, std::string* err);

// gen.cpp
u64 splitmix64_next(u64* state);
int gen_er(int32_t n, int64_t m_undirected, u64 seed, HostGraph* g);
int gen_rmat(int scale, int edge_factor, double a, double b, double c, u64 seed, HostGraph* g);
int gen_chung_lu(int32_t n, double gamma, double mean_deg, double max_deg, u64 seed, HostGraph* g);
void gen_random_motifs(u64 seed, int count, int core_min, int core_max, int fringe_max, std::vector<Pattern>* out);
int csr_from_pairs(int32_t n, std::vector<u64>& keys, HostGraph* g);  // keys = (src << 32) | dst

}  // namespace fsgc
