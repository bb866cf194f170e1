// fringePreprocess <pattern#> | fringePreprocess <nodes> <s> <d> <s> <d> ...
//
// Drop-in for the reference's fringePreprocess (src/fringePreprocess.cpp:287-365):
// same arguments, writes ./motif.mo (139 bytes; unused slots zero) and ./motif.dot,
// prints the same report.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/fsgc.h"

int main(int argc, char* argv[])
{
  printf("MI355X Fringe Preprocessor (fsgc)\n\n");
  if ((argc < 2) || ((argc % 2) != 0)) {
    printf("USAGE: %s motif_number\n", argv[0]);
    printf("USAGE: %s number_of_nodes edge1_source edge1_destination ...\n\n", argv[0]);
    return -1;
  }
  uint8_t mo[FSGC_MO_BYTES];
  static char log[1 << 16], dot[1 << 16];
  int rc;
  if (argc == 2) {
    rc = fsgc_preprocess_pattern(atoi(argv[1]), mo, log, sizeof(log), dot, sizeof(dot));
  } else {
    std::vector<int> e;
    for (int p = 2; p + 1 < argc; p += 2) { e.push_back(atoi(argv[p])); e.push_back(atoi(argv[p + 1])); }
    rc = fsgc_preprocess(atoi(argv[1]), (int)e.size() / 2, e.data(), mo, log, sizeof(log), dot, sizeof(dot));
  }
  printf("%s", log);
  if (rc) { fprintf(stderr, "%s\n", fsgc_last_error()); return -1; }
  FILE* f = fopen("motif.dot", "w");
  if (f) { fputs(dot, f); fclose(f); }
  f = fopen("motif.mo", "wb");
  const size_t size = f ? fwrite(mo, 1, sizeof(mo), f) : 0;
  if (f) fclose(f);
  if (size != sizeof(mo)) { fprintf(stderr, "ERROR: could not write matching order to file\n"); return -1; }
  printf("wrote matching order to file 'motif.mo'\n\n");
  return 0;
}
