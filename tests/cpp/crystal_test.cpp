// Addressing laws of nmhd/crystal.h, the checks of the reference's
// test/crystal/basic_test.curb:45-111 on a 20 x 20 x 10 lattice
// (test/crystal/simulation_config.hrb): address <-> coordinate round trip,
// whole steps move 2 and half steps move 1 in the doubled lattice (point with
// lattice bits b at index i sits at 2*i + b_d), and half steps flip the bit.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nmhd/crystal.h"

using namespace nmhd;

static int failures = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      ++failures;                                                  \
      if (failures < 20) std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
    }                                                              \
  } while (0)

static Lattice L = {{20, 20, 10}, {0.05f, 0.05f, 0.1f}};

template <int Bit>
Triplet<int> big(const RCo<Bit> &r) {
  return Triplet<int>(2 * r[X] + (Bit & 1), 2 * r[Y] + ((Bit >> 1) & 1), 2 * r[Z] + ((Bit >> 2) & 1));
}

template <int Bit>
void check_bit(long i) {
  const RCo<Bit> r(L, i);
  EXPECT(r.addr() == i);
  EXPECT(big(r) == big(r + X + Y + Z - X - Y - Z));
  EXPECT(big(r + X) == big(r) + Triplet<int>(2, 0, 0));
  EXPECT(big(r + Y) == big(r) + Triplet<int>(0, 2, 0));
  EXPECT(big(r + Z) == big(r) + Triplet<int>(0, 0, 2));
  EXPECT(big(r - X) == big(r) - Triplet<int>(2, 0, 0));
  EXPECT(big(r - Y) == big(r) - Triplet<int>(0, 2, 0));
  EXPECT(big(r - Z) == big(r) - Triplet<int>(0, 0, 2));
  EXPECT(big(r + HX) == big(r) + Triplet<int>(1, 0, 0));
  EXPECT(big(r + HY) == big(r) + Triplet<int>(0, 1, 0));
  EXPECT(big(r + HZ) == big(r) + Triplet<int>(0, 0, 1));
  EXPECT(big(r - HX) == big(r) - Triplet<int>(1, 0, 0));
  EXPECT(big(r - HY) == big(r) - Triplet<int>(0, 1, 0));
  EXPECT(big(r - HZ) == big(r) - Triplet<int>(0, 0, 1));
  EXPECT((r + HX).bit() == (Bit ^ 1) && (r - HY).bit() == (Bit ^ 2) && (r + HZ).bit() == (Bit ^ 4));
}

int main() {
  const long n = (long)L.n[0] * L.n[1] * L.n[2];
  std::srand(12345);
  int done = 0;
  while (done < 300) {
    const long i = std::rand() % n;
    const RCo<0> r(L, i);
    if (r[X] <= 0 || r[X] >= L.n[0] - 1 || r[Y] <= 0 || r[Y] >= L.n[1] - 1 || r[Z] <= 0 ||
        r[Z] >= L.n[2] - 1)
      continue;  // interior points only, as the reference test does
    check_bit<0>(i); check_bit<1>(i); check_bit<2>(i); check_bit<3>(i);
    check_bit<4>(i); check_bit<5>(i); check_bit<6>(i); check_bit<7>(i);
    ++done;
  }
  // periodic wrap of addresses and the +0.5 position of every point
  const RCo<7> c(L, 0, 0, 0);
  EXPECT((c - X).addr() == L.n[0] - 1);
  EXPECT((c - HX).addr() == 0 && (c - HX - HX).addr() == L.n[0] - 1);
  const Triplet<Real> p = (c - HX).position(Triplet<Real>(0, 0, 0));
  EXPECT(p[0] == 0.5f * L.d[0] && p[1] == 0.5f * L.d[1]);
  // Mesh views index by address and re-type with shift()
  std::vector<Real> buf(n);
  for (long i = 0; i < n; ++i) buf[i] = Real(i);
  BodyScalar s(&buf[0]);
  const RCo<7> q(L, 3, 4, 5);
  EXPECT(s[q] == Real(q.addr()));
  Mesh<6, Real> f = s.shift(X);
  EXPECT(f[q - HX] == s[q]);
  std::printf("crystal_test: %d failures\n", failures);
  return failures ? 1 : 0;
}
