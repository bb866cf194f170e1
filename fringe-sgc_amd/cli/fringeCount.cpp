// fringeCount <graph.egr> <motif.mo> [--device N] [--slice MOD REM] [--relabel asc|desc|none]
//
// Drop-in for the reference's fringeCount (src/fringeCount.cu:2265-2331): same
// arguments, the same "runtime: %.6f s" and "%lld occurrences" lines (the u64 printed
// signed, as the reference does), exit status 255 on errors. One extra line follows
// with the exact unsigned 128-bit count and the overflow flags.
//
// Like the reference, which renumbers the graph by degree inside its timed region
// before a triangle-core count (fringeCount.cu:1348), wedge-core motifs are counted on
// the graph renumbered by ascending degree (fsgc_graph_relabel; the pivot kernel runs
// in 2.24 s on R-MAT s20 against 3.47 s on the file's ids); the relabelling is inside the
// "runtime" line. --relabel overrides the choice; counts do not depend on it. Root
// slices (--slice) are defined on the file's ids and are never relabelled.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/fsgc.h"

static std::string u128_str(uint64_t lo, uint64_t hi)
{
  unsigned __int128 x = ((unsigned __int128)hi << 64) | lo;
  if (x == 0) return "0";
  std::string s;
  while (x) { s.insert(s.begin(), (char)('0' + (int)(x % 10))); x /= 10; }
  return s;
}

int main(int argc, char* argv[])
{
  printf("MI355X Fringe Counting (fsgc, HIP/gfx950)\n\n");
  if (argc < 3) {
    printf("USAGE: %s input_graph motif_matching_order\n", argv[0]);
    return -1;
  }
  int device = 0;
  int relabel = -1;  // -1: by motif class
  fsgc_tune tune;
  memset(&tune, 0, sizeof(tune));
  for (int i = 3; i < argc; i++) {
    if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--slice") && i + 2 < argc) { tune.slice_mod = atoi(argv[++i]); tune.slice_rem = atoi(argv[++i]); }
    else if (!strcmp(argv[i], "--relabel") && i + 1 < argc) {
      const char* o = argv[++i];
      if (!strcmp(o, "asc")) relabel = FSGC_RELABEL_DEGREE_ASC;
      else if (!strcmp(o, "desc")) relabel = FSGC_RELABEL_DEGREE_DESC;
      else if (!strcmp(o, "none")) relabel = -2;
      else { printf("USAGE: %s input_graph motif_matching_order\n", argv[0]); return -1; }
    }
    else { printf("USAGE: %s input_graph motif_matching_order\n", argv[0]); return -1; }
  }
  if (fsgc_init(device)) { fprintf(stderr, "ERROR: %s\n\n", fsgc_last_error()); return -1; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    printf("GPU: %s with %d CUs (%s)\n\n", prop.name, prop.multiProcessorCount, prop.gcnArchName);

  fsgc_graph* g = nullptr;
  if (fsgc_graph_load_egr(argv[1], &g)) { fprintf(stderr, "%s\n\n", fsgc_last_error()); return -1; }
  int32_t n = 0, m = 0, maxdeg = 0;
  fsgc_graph_info(g, &n, &m, &maxdeg);
  printf("input: %s\n", argv[1]);
  printf("nodes: %d\n", n);
  printf("edges: %d (%d)\n\n", m / 2, m);

  uint8_t mo[FSGC_MO_BYTES];
  FILE* f = fopen(argv[2], "rb");
  if (!f) { fprintf(stderr, "ERROR: could not open matching order file %s\n", argv[2]); fsgc_graph_free(g); return -1; }
  const size_t size = fread(mo, 1, sizeof(mo), f);
  fclose(f);
  if (size != sizeof(mo)) { fprintf(stderr, "ERROR: could not read matching order from file\n"); fsgc_graph_free(g); return -1; }
  static char text[8192];
  fsgc_print_matching_order(mo, text, sizeof(text));
  printf("%s", text);

  // wedge core: 3 core vertices, the third adjacent to one earlier core vertex only
  const bool wedge = mo[0] == 3 && __builtin_popcount(mo[9]) == 1;
  if (relabel == -1) relabel = (wedge && tune.slice_mod <= 1) ? FSGC_RELABEL_DEGREE_ASC : -2;
  if (relabel >= 0 && tune.slice_mod > 1) { fprintf(stderr, "ERROR: --slice counts the file's ids; drop --relabel\n\n"); fsgc_graph_free(g); return -1; }

  fsgc_result r;
  const auto t0 = std::chrono::steady_clock::now();
  fsgc_graph* gc = g;
  if (relabel >= 0 && fsgc_graph_relabel(g, relabel, &gc, nullptr, nullptr)) {
    fprintf(stderr, "ERROR: %s\n\n", fsgc_last_error());
    fsgc_graph_free(g);
    return -1;
  }
  const int rc = fsgc_count(gc, mo, &tune, &r);
  const double runtime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (gc != g) fsgc_graph_free(gc);
  if (rc) { fprintf(stderr, "ERROR: %s\n\n", fsgc_last_error()); fsgc_graph_free(g); return -1; }
  static const char* names[] = {"vertex core", "edge core", "wedge core", "triangle core", "general core"};
  printf("%s\n", names[r.motif_class]);
  if (r.motif_class >= FSGC_CLASS_WEDGE) printf("automorphisms: %d\n", (int)(int64_t)r.aut);
  printf("runtime: %.6f s\n", runtime);
  printf("%lld occurrences\n\n", (long long)r.ref_compat);
  printf("exact: %s occurrences (flags 0x%x%s%s%s), %llu core matches, kernel %.3f ms\n",
         u128_str(r.exact_lo, r.exact_hi).c_str(), r.flags, (r.flags & FSGC_OVF64) ? " OVF64" : "",
         (r.flags & FSGC_OVF128) ? " OVF128" : "", (r.flags & FSGC_REF_NCK_INEXACT) ? " REF_NCK_INEXACT" : "",
         (unsigned long long)r.core_matches, r.kernel_ms);
  fsgc_graph_free(g);
  return 0;
}
