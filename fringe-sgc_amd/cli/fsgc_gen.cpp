// fsgc_gen — synthetic inputs and graph facts.
//   fsgc_gen graph <spec> <out.egr>      spec as in fsgc_generate (include/fsgc.h)
//   fsgc_gen motifs <seed> <count> <dir> writes <dir>/motif_XX.mo
//   fsgc_gen info <graph.egr>            nodes, edges, average and maximum degree
//                                        (the reference's src/info.cpp:43-54)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../csrc/fsgc_host.h"
#include "../../include/fsgc.h"

int main(int argc, char* argv[])
{
  if (argc >= 4 && !strcmp(argv[1], "graph")) {
    int32_t n, m, *idx, *adj;
    if (fsgc_generate(argv[2], &n, &m, &idx, &adj)) { fprintf(stderr, "%s\n", fsgc_last_error()); return 1; }
    fsgc::HostGraph g;
    g.n = n;
    g.m = m;
    g.nindex.assign(idx, idx + n + 1);
    g.nlist.assign(adj, adj + m);
    fsgc_free_host(idx);
    fsgc_free_host(adj);
    std::string err;
    if (fsgc::write_egr(argv[3], g, &err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    printf("%s: %d nodes, %d directed edges\n", argv[3], n, m);
    return 0;
  }
  if (argc >= 5 && !strcmp(argv[1], "motifs")) {
    const int count = atoi(argv[3]);
    std::vector<uint8_t> mos((size_t)count * FSGC_MO_BYTES);
    if (fsgc_random_motifs(strtoull(argv[2], nullptr, 10), count, mos.data())) { fprintf(stderr, "%s\n", fsgc_last_error()); return 1; }
    for (int i = 0; i < count; i++) {
      const std::string p = std::string(argv[4]) + "/motif_" + (i < 10 ? "0" : "") + std::to_string(i) + ".mo";
      FILE* f = fopen(p.c_str(), "wb");
      if (!f || fwrite(mos.data() + (size_t)i * FSGC_MO_BYTES, 1, FSGC_MO_BYTES, f) != FSGC_MO_BYTES) { fprintf(stderr, "cannot write %s\n", p.c_str()); return 1; }
      fclose(f);
    }
    return 0;
  }
  if (argc >= 3 && !strcmp(argv[1], "info")) {
    fsgc::HostGraph g;
    std::string err;
    if (fsgc::read_egr(argv[2], &g, &err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    int maxdeg = 0;
    for (int u = 0; u < g.n; u++) maxdeg = std::max(maxdeg, g.nindex[u + 1] - g.nindex[u]);
    printf("nodes: %d\nedges: %lld\navg degree: %.2f\nmax degree: %d\n", g.n, (long long)g.m, (double)g.m / g.n, maxdeg);
    const int rc = fsgc::validate_csr(g, &err);
    printf("valid: %s\n", rc ? err.c_str() : "yes");
    return 0;
  }
  fprintf(stderr, "usage: %s graph <spec> <out.egr> | motifs <seed> <count> <dir> | info <graph.egr>\n", argv[0]);
  return 1;
}
