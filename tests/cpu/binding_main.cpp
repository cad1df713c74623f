// Driver of the reference-side binding (INTEGRATION.md's aline_accp.cpp,
// which tests/test_integration_binding.py extracts into the build
// directory): loads a lattice file with the reference's reader (ref_load of
// oracle/_ref), converts every element with the binding's fill() and writes
// the accp_elem records, each WigTable's table copy after its record, to a
// binary file.  Usage: binding_main lattice.lat out.bin rad_flag
#include <cstdio>
#include <cstdlib>

#include "aline_accp.cpp"

extern "C" void *ref_load(const char *path);

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  ALine *line = static_cast<ALine *>(ref_load(argv[1]));
  if (!line) return 3;
  FILE *f = std::fopen(argv[2], "wb");
  if (!f) return 4;
  const unsigned rad = (unsigned)std::atoi(argv[3]);
  for (unsigned i = 0; i < line->getnumber(); ++i) {
    accp_elem a;
    fill(line->getelement(i), a, rad, true, true, true);
    std::fwrite(&a, sizeof a, 1, f);
    if (a.type == ACCP_WIGTABLE)
      std::fwrite(a.wt_table, sizeof(double), (size_t)(a.wt_mx + a.wt_my + 2 * a.wt_mx * a.wt_my), f);
  }
  std::fclose(f);
  std::printf("%u elements\n", line->getnumber());
  return 0;
}
