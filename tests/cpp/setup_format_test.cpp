// CPU known-answer test of the setup text that save_state / dump write and
// load_state reads (include/nmhd/mhd_solver.h write_setup / read_setup):
//  1. a fixed GlobalSetup + LocalSetup must produce tests/golden/setup_kat.txt
//     byte for byte.  That file was written by hand from the reference's
//     encoder rules (library/mhd_solver.h:110-231 and 437-462,
//     nushio_script/persistent_pod.hrb:80-120): a member count, then one
//     "key value" line per member in declaration order; numbers through a
//     default std::ostream (6 significant digits, %g style: 12.3456789 ->
//     12.3457, 1.5e-7 -> 1.5e-07); strings byte by byte, printable ASCII
//     33..126 except '%' as is, every other byte as %xx (lower-case hex);
//     save_state's layout (INL:1250-1259): GlobalSetup, an empty line,
//     16 '#', LocalSetup, an empty line.
//  2. read_setup of tests/golden/setup_reference_written.txt (members in
//     another order, an unknown member, escapes in output_directory, written
//     by hand in the reference's format) restores exactly the members the
//     reference's readers restore (generation, current_time; mpi_rank,
//     output_directory, mpi_prefix, relative_displacement, relative_velocity)
//     and leaves the others as they were.
// Usage: setup_format_test <golden dir>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "nmhd/mhd_solver.h"

using namespace nmhd;

static int failures = 0;
#define EXPECT(c)                                             \
  do {                                                        \
    if (!(c)) {                                               \
      ++failures;                                             \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
    }                                                         \
  } while (0)

static std::string slurp(const std::string &f) {
  std::ifstream i(f.c_str(), std::ios::binary);
  std::ostringstream o;
  o << i.rdbuf();
  return o.str();
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden";
  // 1. the 256^3 shearing box's setup at generation 123
  GlobalSetup g;
  g.generation_ = 123;
  g.current_time_ = 12.3456789;
  for (int d = 0; d < 3; ++d) {
    g.global_size_[d] = 256;
    g.global_extent_[d] = 1.0f;
    g.d_[d] = 1.0f / 256.0f;
  }
  g.global_corner_[0] = -0.5f;
  g.global_corner_[1] = -0.5f;
  LocalSetup l;
  l.output_directory_ = "run dir/100%";
  l.relative_displacement_ = 1.5e-7;
  l.relative_velocity_ = -1.5;
  l.computational_size_ = 256u * 264u * 256u;
  const int size[3] = {256, 264, 256}, margin[3] = {0, 4, 0};
  for (int d = 0; d < 3; ++d) {
    l.computational_size_d_[d] = size[d];
    l.margin_[d] = margin[d];
    l.d_[d] = 1.0f / 256.0f;
    l.computational_extent_[d] = Real(size[d]) * l.d_[d];
    l.computational_corner_[d] = g.global_corner_[d] + Real(-margin[d]) * l.d_[d];
  }
  std::ostringstream text;
  write_setup(text, g, l);
  const std::string want = slurp(dir + "/setup_kat.txt");
  EXPECT(!want.empty());
  EXPECT(text.str() == want);
  if (text.str() != want) std::cout << "--- written:\n" << text.str() << "--- expected:\n" << want;

  // 2. a setup in the reference's format, written by hand
  GlobalSetup g2;
  LocalSetup l2;
  l2.computational_size_ = 77;
  std::ifstream in((dir + "/setup_reference_written.txt").c_str());
  EXPECT(in.good());
  read_setup(in, g2, l2);
  EXPECT(g2.generation_ == 4096);
  EXPECT(g2.current_time_ == 3.14159);
  EXPECT(g2.global_size_[0] == 0);  // not restored by the reference either
  EXPECT(l2.mpi_rank_ == 3);
  EXPECT(l2.output_directory_ == "/data/mri runs/q=1.5%/\t");
  EXPECT(l2.mpi_prefix_ == "0003");
  EXPECT(l2.relative_displacement_ == 0.25);
  EXPECT(l2.relative_velocity_ == -1.5);
  EXPECT(l2.computational_size_ == 77u);
  // the restored string re-encodes to the escaped form it was read from
  EXPECT(pod::enc(l2.output_directory_) == "/data/mri%20runs/q=1.5%25/%09");
  std::printf("setup_format_test: %d failures\n", failures);
  return failures ? 1 : 0;
}
