// nmhd/crystal.h -- staggered-mesh addressing types of the C++ facade.
//
// Host-side counterparts of the reference's axis tags, Triplet, RCo and Mesh
// (library/direction.h:3-234, library/crystal.h:13-131), with the box size
// a run-time value instead of the compile-time kSizeX/Y/Z:
//  - RCo<Bit> is a coordinate in the staggered lattice; bit d of Bit is 1
//    when the point is cell-centred along axis d and 0 on the lower d-face.
//  - r + X / r - X move one cell; r + HX / r - HX move half a cell and flip
//    the bit: cell i - HX is its lower face i, cell i + HX its upper face i+1,
//    face i + HX is cell i, face i - HX is cell i-1 (crystal.h:191-663).
//  - addresses wrap periodically on every axis, (i + N) % N
//    (CRYSTAL_PERIODIC_*, crystal.h:31-50); addr = (z*ny + y)*nx + x.
//  - position() reproduces the reference's positions, which put every
//    staggered point at the centre of its index because `Bit & 1 == 0`
//    parses as `Bit & (1 == 0)` (crystal.h:78-106).
#ifndef NMHD_CRYSTAL_H
#define NMHD_CRYSTAL_H

namespace nmhd {

typedef float Real;
typedef double PreciseReal;

struct XHalf;
struct YHalf;
struct ZHalf;
struct YAxis;
struct ZAxis;
struct XAxis {
  enum { mask = 1, index = 0 };
  typedef XHalf Half;
  typedef ZAxis Prev;
  typedef YAxis Next;
};
struct YAxis {
  enum { mask = 2, index = 1 };
  typedef YHalf Half;
  typedef XAxis Prev;
  typedef ZAxis Next;
};
struct ZAxis {
  enum { mask = 4, index = 2 };
  typedef ZHalf Half;
  typedef YAxis Prev;
  typedef XAxis Next;
};
struct XHalf {
  enum { mask = 1, index = 0 };
  typedef XAxis Whole;
};
struct YHalf {
  enum { mask = 2, index = 1 };
  typedef YAxis Whole;
};
struct ZHalf {
  enum { mask = 4, index = 2 };
  typedef ZAxis Whole;
};
static const XAxis X = XAxis();
static const YAxis Y = YAxis();
static const ZAxis Z = ZAxis();
static const XHalf HX = XHalf();
static const YHalf HY = YHalf();
static const ZHalf HZ = ZHalf();

template <class T>
struct Triplet {
  T x_, y_, z_;
  Triplet() : x_(), y_(), z_() {}
  Triplet(const T &x, const T &y, const T &z) : x_(x), y_(y), z_(z) {}
  T &operator[](int d) { return d == 0 ? x_ : (d == 1 ? y_ : z_); }
  const T &operator[](int d) const { return d == 0 ? x_ : (d == 1 ? y_ : z_); }
  T &operator[](XAxis) { return x_; }
  T &operator[](YAxis) { return y_; }
  T &operator[](ZAxis) { return z_; }
  const T &operator[](XAxis) const { return x_; }
  const T &operator[](YAxis) const { return y_; }
  const T &operator[](ZAxis) const { return z_; }
  Triplet operator+(const Triplet &o) const { return Triplet(x_ + o.x_, y_ + o.y_, z_ + o.z_); }
  Triplet operator-(const Triplet &o) const { return Triplet(x_ - o.x_, y_ - o.y_, z_ - o.z_); }
  bool operator==(const Triplet &o) const { return x_ == o.x_ && y_ == o.y_ && z_ == o.z_; }
  bool operator!=(const Triplet &o) const { return !(*this == o); }
};

// Array size (ghosts included) and spacing of a box.
struct Lattice {
  int n[3];
  Real d[3];
};

template <int Bit>
struct RCo {
  const Lattice *lat;
  int v[3];  // unwrapped indices
  RCo(const Lattice &l, int x, int y, int z) : lat(&l) {
    v[0] = x;
    v[1] = y;
    v[2] = z;
  }
  RCo(const Lattice &l, long addr) : lat(&l) {
    v[0] = (int)(addr % l.n[0]);
    v[1] = (int)((addr / l.n[0]) % l.n[1]);
    v[2] = (int)(addr / l.n[0] / l.n[1]);
  }
  int bit() const { return Bit; }
  int wrapped(int d) const { return ((v[d] % lat->n[d]) + lat->n[d]) % lat->n[d]; }
  int operator[](XAxis) const { return wrapped(0); }
  int operator[](YAxis) const { return wrapped(1); }
  int operator[](ZAxis) const { return wrapped(2); }
  long addr() const {
    return ((long)wrapped(2) * lat->n[1] + wrapped(1)) * lat->n[0] + wrapped(0);
  }
  Triplet<Real> position(const Triplet<Real> &corner) const {
    return Triplet<Real>(corner[0] + lat->d[0] * (Real(wrapped(0)) + Real(0.5f)),
                         corner[1] + lat->d[1] * (Real(wrapped(1)) + Real(0.5f)),
                         corner[2] + lat->d[2] * (Real(wrapped(2)) + Real(0.5f)));
  }
};

// whole steps along an axis keep the lattice type
template <int Bit, int D>
RCo<Bit> step_whole(RCo<Bit> r, int s) {
  r.v[D] += s;
  return r;
}
// half steps flip the bit: from a cell (bit set) +H reaches its upper face
// (index + 1) and -H its lower face (same index); from a face +H reaches the
// cell above (same index) and -H the cell below (index - 1)
template <int Bit, int D, int M>
RCo<Bit ^ M> step_half(const RCo<Bit> &r, int s) {
  RCo<Bit ^ M> o(*r.lat, r.v[0], r.v[1], r.v[2]);
  const bool cell = (Bit & M) != 0;
  if (cell && s > 0) o.v[D] += 1;
  if (!cell && s < 0) o.v[D] -= 1;
  return o;
}
template <int Bit> RCo<Bit> operator+(const RCo<Bit> &r, XAxis) { return step_whole<Bit, 0>(r, 1); }
template <int Bit> RCo<Bit> operator+(const RCo<Bit> &r, YAxis) { return step_whole<Bit, 1>(r, 1); }
template <int Bit> RCo<Bit> operator+(const RCo<Bit> &r, ZAxis) { return step_whole<Bit, 2>(r, 1); }
template <int Bit> RCo<Bit> operator-(const RCo<Bit> &r, XAxis) { return step_whole<Bit, 0>(r, -1); }
template <int Bit> RCo<Bit> operator-(const RCo<Bit> &r, YAxis) { return step_whole<Bit, 1>(r, -1); }
template <int Bit> RCo<Bit> operator-(const RCo<Bit> &r, ZAxis) { return step_whole<Bit, 2>(r, -1); }
template <int Bit> RCo<Bit ^ 1> operator+(const RCo<Bit> &r, XHalf) { return step_half<Bit, 0, 1>(r, 1); }
template <int Bit> RCo<Bit ^ 2> operator+(const RCo<Bit> &r, YHalf) { return step_half<Bit, 1, 2>(r, 1); }
template <int Bit> RCo<Bit ^ 4> operator+(const RCo<Bit> &r, ZHalf) { return step_half<Bit, 2, 4>(r, 1); }
template <int Bit> RCo<Bit ^ 1> operator-(const RCo<Bit> &r, XHalf) { return step_half<Bit, 0, 1>(r, -1); }
template <int Bit> RCo<Bit ^ 2> operator-(const RCo<Bit> &r, YHalf) { return step_half<Bit, 1, 2>(r, -1); }
template <int Bit> RCo<Bit ^ 4> operator-(const RCo<Bit> &r, ZHalf) { return step_half<Bit, 2, 4>(r, -1); }

template <int Bit, class T>
struct Mesh {
  T *ptr_;
  explicit Mesh(T *p = 0) : ptr_(p) {}
  T &operator[](const RCo<Bit> &r) const { return ptr_[r.addr()]; }
  template <class Axis>
  Mesh<Bit ^ Axis::mask, T> shift(Axis) const {
    return Mesh<Bit ^ Axis::mask, T>(ptr_);
  }
  T *ptr() const { return ptr_; }
};

// Field views of the state (library/physics_util.h:4-7): cell-centred
// scalars and vectors on Bit 7, face B_d on the lower d-face (Bit 7 ^ mask).
typedef Mesh<7, Real> BodyScalar;
typedef Triplet<Mesh<7, Real> > BodyVector;
struct FaceVector {
  Mesh<6, Real> x_;
  Mesh<5, Real> y_;
  Mesh<3, Real> z_;
  FaceVector(Real *bx = 0, Real *by = 0, Real *bz = 0) : x_(bx), y_(by), z_(bz) {}
  Mesh<6, Real> &operator[](XAxis) { return x_; }
  Mesh<5, Real> &operator[](YAxis) { return y_; }
  Mesh<3, Real> &operator[](ZAxis) { return z_; }
};

}  // namespace nmhd

#endif  // NMHD_CRYSTAL_H
