/*
 * LATfield2_mi.hpp -- C++ mirror of the LATfield2 API for the particle-mesh path,
 * running on one MI355X through the C ABI of lf2mi.h.
 *
 * Same names, argument order and meaning as the reference, so that a driver written for
 * LATfield2 compiles against this header (include "LATfield2.hpp" from this directory, or
 * this file) and links with -llf2mi:
 *   Parallel2d, parallel, COUT   LATfield2_parallel2d_decl.hpp:20-500 (one rank)
 *   Lattice                      LATfield2_Lattice_decl.hpp:13-218
 *   Site, rKSite, cKSite         LATfield2_Site_decl.hpp:13-277 (rKSite dims {0,2,1},
 *                                cKSite dims {1,2,0}), halo iteration :99-119
 *   forallboundary_start/_stop   looping_macro.hpp:10-21
 *   Imag                         Imag.hpp:24-110
 *   Field<Real>, Field<Imag>     LATfield2_Field_decl.hpp:61-430 (AoS; matrix fields with
 *                                the reference's component order :760-807)
 *   Field::updateHalo            LATfield2_Field_decl.hpp:889-935
 *   PlanFFT<Imag>, execute       LATfield2_PlanFFT_decl.hpp:83-1203 (R2C and C2C,
 *                                FFT_IN_PLACE aliasing :804-824)
 *   Particles<part,info,dtype>   particles/LATfield2_Particles.hpp:131-344, updateVel and
 *                                moveParticles with output[]/reduce_type[] (:217-244)
 *   projection_init, scalar/vector/VecVec ProjectionCIC _project/_comm and the CIC-NGP
 *   hybrids                      particles/projections.hpp:31-1814
 *
 * Memory: every Field owns a device buffer (the data the kernels use) and a host mirror
 * (data_, what data() and operator() return, as the reference's host array).  The two are
 * kept coherent automatically: a host access (operator(), data()) first brings the host
 * copy up to date if the device was written since, and marks it as possibly written; a
 * device operation (updateHalo, PlanFFT::execute, the projections, updateVel) first uploads
 * a host copy that was touched.  A reference driver therefore needs no extra calls.
 * toDevice()/toHost() force the transfers; device_ptr() hands the device buffer to
 * callers that launch their own kernels.  Writes through the public data_ pointer are not
 * seen: use data() or operator(), or call touch_host() afterwards.  The host mirror is
 * allocated on the first host access.
 *
 * updateVel / moveParticles accept functors with the reference's signature.  The ones
 * with a device implementation are updateVel_simple (updateVel_function.hpp:13-38, the
 * identity), updateVel_cicKick (the PM gather + kick: v += dtau * CIC(-grad phi) of a
 * potential fields[0]) and move_particles_simple (move_function.hpp:15-32, the drift
 * pos += dtau vel with the periodic wrap and re-binning of Particles.hpp:2028-2056).
 * Any other functor is refused with a message and exit, as the reference aborts.  Their
 * per-particle outputs are, by slot, |v|^2, m v_x, m v_y, m v_z, m, m |v|^2 / 2 (after
 * the update), reduced by reduce_type (SUM, MIN, MAX and the _LOCAL forms, which are the
 * same on one rank) in FP64 on the device.
 *
 * Errors print the library's message and exit(-1), as the reference aborts.
 * Precision: Real is double, or float when SINGLE is defined, as in Imag.hpp:9-15.
 * Box length 1 (particles/LATfield2_Particles.hpp:464-467 bins with it).
 */
#ifndef LATFIELD2_MI_HPP
#define LATFIELD2_MI_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <forward_list>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "lf2mi.h"

// The reference's LATfield2.hpp:63 opens namespace std for its drivers (cout, endl, max,
// string ...), and they rely on it.
using namespace std;

// reduction types of the updateVel / moveParticles outputs (particles/LATfield2_Particles.hpp:81-88)
#ifndef SUM
#define SUM 1
#define MIN 2
#define MAX 4
#define SUM_LOCAL 8
#define MIN_LOCAL 16
#define MAX_LOCAL 32
#endif
#ifndef MAX_NUMBER
#define MAX_NUMBER 9223372036854775807
#endif

// The timer and barrier the reference drivers take from mpi.h (benchmarks_tests/wave_test.cpp
// times with MPI_Wtime), for one process.  Left out when a real mpi.h was included first.
#ifndef MPI_VERSION
typedef int MPI_Comm;
#define MPI_COMM_WORLD 0
inline double MPI_Wtime() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline int MPI_Barrier(MPI_Comm) { return 0; }
#endif

// ------------------------------------------------------------------ Parallel2d
// The process grid of LATfield2_parallel2d.hpp:42-225 on one rank: initialize(n, m) accepts
// only n = m = 1 (the reference aborts when n*m differs from the number of processes), and
// every collective is the identity (a reduction or broadcast over one rank returns its
// input, LATfield2_parallel2d.hpp:266-314).  barrier() also waits for the device, so host
// timers around it measure finished work.
class Parallel2d {
  public:
    Parallel2d() {
        grid_size_[0] = grid_size_[1] = 1;
        grid_rank_[0] = grid_rank_[1] = 0;
        last_proc_[0] = last_proc_[1] = true;
    }
    void initialize(int proc_size0, int proc_size1) { initialize(proc_size0, proc_size1, 0, 0); }
    void initialize(int proc_size0, int proc_size1, int IO_total_size, int IO_node_size) {
        (void)IO_node_size;
        if (proc_size0 * proc_size1 != 1 || IO_total_size != 0) {
            cerr << "Latfield::Parallel2d::initialization - wrong number of process" << endl;
            cerr << "Latfield::Parallel2d::initialization - Number of total process must be equal to "
                    "proc_size0*proc_size1 (one process drives the GPU here: use -n 1 -m 1)"
                 << endl;
            abortForce();
        }
        grid_size_[0] = proc_size0;
        grid_size_[1] = proc_size1;
        initialized_ = true;
    }
    void abortForce() { std::exit(EXIT_FAILURE); }
    void abortRequest() { std::exit(EXIT_FAILURE); }
    void barrier() { (void)lf2mi_synchronize(); }

    template <class Type> void broadcast(Type &, int) {}
    template <class Type> void broadcast(Type *, int, int) {}
    template <class Type> void broadcast_dim0(Type &, int) {}
    template <class Type> void broadcast_dim0(Type *, int, int) {}
    template <class Type> void broadcast_dim1(Type &, int) {}
    template <class Type> void broadcast_dim1(Type *, int, int) {}

    template <class Type> void sum(Type &) {}
    template <class Type> void sum(Type *, int) {}
    template <class Type> void sum_dim0(Type &) {}
    template <class Type> void sum_dim0(Type *, int) {}
    template <class Type> void sum_dim1(Type &) {}
    template <class Type> void sum_dim1(Type *, int) {}
    template <class Type> void sum_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void sum_to(Type *, int, int dest = 0) { check_dest(dest); }
    template <class Type> void sum_dim0_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void sum_dim0_to(Type *, int, int dest = 0) { check_dest(dest); }
    template <class Type> void sum_dim1_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void sum_dim1_to(Type *, int, int dest = 0) { check_dest(dest); }

    template <class Type> void max(Type &) {}
    template <class Type> void max(Type *, int) {}
    template <class Type> void max_dim0(Type &) {}
    template <class Type> void max_dim0(Type *, int) {}
    template <class Type> void max_dim1(Type &) {}
    template <class Type> void max_dim1(Type *, int) {}
    template <class Type> void max_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void max_to(Type *, int, int dest = 0) { check_dest(dest); }
    template <class Type> void max_dim0_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void max_dim0_to(Type *, int, int dest = 0) { check_dest(dest); }
    template <class Type> void max_dim1_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void max_dim1_to(Type *, int, int dest = 0) { check_dest(dest); }

    template <class Type> void min(Type &) {}
    template <class Type> void min(Type *, int) {}
    template <class Type> void min_dim0(Type &) {}
    template <class Type> void min_dim0(Type *, int) {}
    template <class Type> void min_dim1(Type &) {}
    template <class Type> void min_dim1(Type *, int) {}
    template <class Type> void min_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void min_to(Type *, int, int dest = 0) { check_dest(dest); }
    template <class Type> void min_dim0_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void min_dim0_to(Type *, int, int dest = 0) { check_dest(dest); }
    template <class Type> void min_dim1_to(Type &, int dest = 0) { check_dest(dest); }
    template <class Type> void min_dim1_to(Type *, int, int dest = 0) { check_dest(dest); }

    // messages to oneself: send queues a copy, receive takes the oldest one
    template <class Type> void send(Type &message, int to) { send(&message, 1, to); }
    template <class Type> void send(Type *array, int len, int to) {
        check_dest(to);
        const char *b = reinterpret_cast<const char *>(array);
        mailbox_.emplace_back(b, b + sizeof(Type) * (size_t)len);
    }
    template <class Type> void send_dim0(Type &message, int to) { send(&message, 1, to); }
    template <class Type> void send_dim0(Type *array, int len, int to) { send(array, len, to); }
    template <class Type> void send_dim1(Type &message, int to) { send(&message, 1, to); }
    template <class Type> void send_dim1(Type *array, int len, int to) { send(array, len, to); }
    template <class Type> void receive(Type &message, int from) { receive(&message, 1, from); }
    template <class Type> void receive(Type *array, int len, int from) {
        check_dest(from);
        if (mailbox_.empty() || mailbox_.front().size() != sizeof(Type) * (size_t)len) {
            cerr << "Latfield::Parallel2d::receive - no matching message (one process)" << endl;
            abortForce();
        }
        std::memcpy(array, mailbox_.front().data(), mailbox_.front().size());
        mailbox_.pop_front();
    }
    template <class Type> void receive_dim0(Type &message, int from) { receive(&message, 1, from); }
    template <class Type> void receive_dim0(Type *array, int len, int from) { receive(array, len, from); }
    template <class Type> void receive_dim1(Type &message, int from) { receive(&message, 1, from); }
    template <class Type> void receive_dim1(Type *array, int len, int from) { receive(array, len, from); }

    // the neighbour exchanges of the rod decomposition: on one rank the neighbour in both
    // directions is this process, so what is sent up (down) arrives from below (above)
    template <class Type> void sendUp_dim0(Type &bufferSend, Type &bufferRec, long len) { copy_buf(bufferSend, bufferRec, len); }
    template <class Type> void sendDown_dim0(Type &bufferSend, Type &bufferRec, long len) { copy_buf(bufferSend, bufferRec, len); }
    template <class Type> void sendUp_dim1(Type &bufferSend, Type &bufferRec, long len) { copy_buf(bufferSend, bufferRec, len); }
    template <class Type> void sendDown_dim1(Type &bufferSend, Type &bufferRec, long len) { copy_buf(bufferSend, bufferRec, len); }
    template <class Type>
    void sendUpDown_dim0(Type &sendUp, Type &recUp, long lenUp, Type &sendDown, Type &recDown, long lenDown) {
        copy_buf(sendUp, recUp, lenUp);
        copy_buf(sendDown, recDown, lenDown);
    }
    template <class Type>
    void sendUpDown_dim1(Type &sendUp, Type &recUp, long lenUp, Type &sendDown, Type &recDown, long lenDown) {
        copy_buf(sendUp, recUp, lenUp);
        copy_buf(sendDown, recDown, lenDown);
    }

    int size() { return 1; }
    int rank() { return 0; }
    int world_size() { return 1; }
    int world_rank() { return 0; }
    int *grid_size() { return grid_size_; }
    int *grid_rank() { return grid_rank_; }
    int root() { return 0; }
    bool isRoot() { return true; }
    bool *last_proc() { return last_proc_; }
    int grid2world(int n, int m) { return n + grid_size_[0] * m; }
    bool isIO() { return false; }
    bool initialized() const { return initialized_; }
    void PleaseNeverFinalizeMPI() {}

  private:
    void check_dest(int r) {
        if (r != 0) {
            cerr << "Latfield::Parallel2d: rank " << r << " does not exist (one process)" << endl;
            abortForce();
        }
    }
    // Type is an array pointer (Real *, ...) as in the reference's fold buffers
    template <class Type> static void copy_buf(Type &src, Type &dst, long len) {
        if (len > 0 && &src[0] != &dst[0]) std::memmove(&dst[0], &src[0], sizeof(src[0]) * (size_t)len);
    }
    int grid_size_[2], grid_rank_[2];
    bool last_proc_[2];
    bool initialized_ = false;
    std::deque<std::vector<char>> mailbox_;
};

// the single instance the reference defines in its header (LATfield2.hpp:78)
inline Parallel2d parallel;

#define COUT if (parallel.isRoot()) cout

namespace LATfield2 {

#ifdef SINGLE
typedef float Real;
#else
typedef double Real;
#endif

const int FFT_FORWARD = 1;
const int FFT_BACKWARD = -1;
const int FFT_IN_PLACE = 16;
const int FFT_OUT_OF_PLACE = -16;
const int symmetric = 1;
const int unsymmetric = 0;
const int FROM_INFO = 0;
const int FROM_PART = 1;

inline int lf2mi_real_code() { return sizeof(Real) == 4 ? LF2MI_F32 : LF2MI_F64; }

inline void lf2mi_or_die(int rc, const char *where) {
    if (rc != LF2MI_OK) {
        std::fprintf(stderr, "LATfield2::%s: %s\nLATfield2 : Abort Process Requested\n", where, lf2mi_last_error());
        std::exit(-1);
    }
}
inline void lf2mi_die(const char *msg) {
    std::fprintf(stderr, "%s\nLATfield2 : Abort Process Requested\n", msg);
    std::exit(-1);
}

// ------------------------------------------------------------------ Imag
class Imag {
  public:
    Imag() : re_(0), im_(0) {}
    Imag(Real re, Real im) : re_(re), im_(im) {}
    Imag operator-() const { return Imag(-re_, -im_); }
    Imag operator+(const Imag &o) const { return Imag(re_ + o.re_, im_ + o.im_); }
    Imag operator-(const Imag &o) const { return Imag(re_ - o.re_, im_ - o.im_); }
    Imag operator*(const Imag &o) const { return Imag(re_ * o.re_ - im_ * o.im_, re_ * o.im_ + im_ * o.re_); }
    Imag operator/(const Imag &o) const {
        const Real d = o.norm();
        return Imag((re_ * o.re_ + im_ * o.im_) / d, (im_ * o.re_ - re_ * o.im_) / d);
    }
    void operator=(Real r) { re_ = r; im_ = 0; }
    friend Imag operator+(const Imag &z, Real a) { return Imag(z.re_ + a, z.im_); }
    friend Imag operator+(Real a, const Imag &z) { return Imag(z.re_ + a, z.im_); }
    friend Imag operator-(const Imag &z, Real a) { return Imag(z.re_ - a, z.im_); }
    friend Imag operator-(Real a, const Imag &z) { return Imag(a - z.re_, z.im_); }
    friend Imag operator*(const Imag &z, Real a) { return Imag(z.re_ * a, z.im_ * a); }
    friend Imag operator*(Real a, const Imag &z) { return Imag(z.re_ * a, z.im_ * a); }
    friend Imag operator/(const Imag &z, Real a) { return Imag(z.re_ / a, z.im_ / a); }
    void operator+=(const Imag &z) { re_ += z.re_; im_ += z.im_; }
    void operator-=(const Imag &z) { re_ -= z.re_; im_ -= z.im_; }
    void operator*=(const Imag &z) {
        const Real r = re_ * z.re_ - im_ * z.im_;
        im_ = re_ * z.im_ + im_ * z.re_;
        re_ = r;
    }
    void operator+=(Real a) { re_ += a; }
    void operator-=(Real a) { re_ -= a; }
    void operator*=(Real a) { re_ *= a; im_ *= a; }
    void operator/=(Real a) { re_ /= a; im_ /= a; }
    Real &real() { return re_; }
    Real &imag() { return im_; }
    Real real() const { return re_; }
    Real imag() const { return im_; }
    Real phase() const { return std::acos(re_ / std::sqrt(re_ * re_ + im_ * im_)) * (im_ < 0 ? -1 : 1); }
    Imag conj() const { return Imag(re_, -im_); }
    Real norm() const { return re_ * re_ + im_ * im_; }
    friend Imag sin(const Imag &z) { return Imag(std::sin(z.re_) * std::cosh(z.im_), std::cos(z.re_) * std::sinh(z.im_)); }
    friend Imag cos(const Imag &z) { return Imag(std::cos(z.re_) * std::cosh(z.im_), -std::sin(z.re_) * std::sinh(z.im_)); }
    friend std::ostream &operator<<(std::ostream &os, const Imag &z) { return os << z.re_ << " + i" << z.im_; }
    friend std::istream &operator>>(std::istream &is, Imag &z) { return is >> z.re_ >> z.im_; }

  private:
    Real re_, im_;
};
inline Imag expi(Real x) { return Imag(std::cos(x), std::sin(x)); }

// ------------------------------------------------------------------ Lattice
class Lattice {
  public:
    Lattice() : dim_(0), halo_(0) {
        for (int i = 0; i < 3; ++i) { size_[i] = 0; jump_[i] = 0; }
        coordSkip_[0] = coordSkip_[1] = 0;
    }
    Lattice(int dim, const int size, int halo) : Lattice() { initialize(dim, size, halo); }
    Lattice(int dim, const int *size, int halo) : Lattice() { initialize(dim, size, halo); }

    void initialize(int dim, const int size, int halo) {
        int s[3] = {size, size, size};
        initialize(dim, s, halo);
    }
    void initialize(int dim, const int *size, int halo) {
        if (dim != 3) lf2mi_die("LATfield2::Lattice: the MI355X PM path supports 3 dimensions");
        dim_ = dim;
        halo_ = halo;
        for (int i = 0; i < 3; ++i) size_[i] = size[i];
        jump_[0] = 1;
        for (int i = 1; i < 3; ++i) jump_[i] = jump_[i - 1] * (size_[i - 1] + 2 * halo_);
    }
    // (N/2+1, N, N) k-lattice of the R2C transform (LATfield2_Lattice.hpp:225-259)
    void initializeRealFFT(Lattice &lat_real, int halo) {
        if (lat_real.dim() != 3 || lat_real.size(0) != lat_real.size(1) || lat_real.size(1) != lat_real.size(2))
            lf2mi_die("Latfield2d::Lattice::initializeRealFFT : fft curently work only for 3d cubic lattice");
        int s[3] = {lat_real.size(0) / 2 + 1, lat_real.size(0), lat_real.size(0)};
        initialize(3, s, halo);
    }
    // N^3 k-lattice of the C2C transform (LATfield2_Lattice.hpp:260-297)
    void initializeComplexFFT(Lattice &lat_real, int halo) {
        if (lat_real.dim() != 3 || lat_real.size(0) != lat_real.size(1) || lat_real.size(1) != lat_real.size(2))
            lf2mi_die("Latfield2d::Lattice::initializeComplexFFT : fft curently work only for 3d cubic lattice");
        initialize(3, lat_real.size(0), halo);
    }

    int dim() const { return dim_; }
    int size(int i) const { return size_[i]; }
    int *size() { return size_; }
    int sizeLocal(int i) const { return size_[i]; }
    int *sizeLocal() { return size_; }
    int halo() const { return halo_; }
    long jump(int i) const { return jump_[i]; }
    long sites() const { return (long)size_[0] * size_[1] * size_[2]; }
    long sitesLocal() const { return sites(); }
    long sitesGross() const { return (long)(size_[0] + 2 * halo_) * (size_[1] + 2 * halo_) * (size_[2] + 2 * halo_); }
    long sitesLocalGross() const { return sitesGross(); }
    long siteFirst() const { return halo_ * (jump_[0] + jump_[1] + jump_[2]); }
    long siteLast() const { return siteFirst() + (size_[0] - 1) * jump_[0] + (size_[1] - 1) * jump_[1] + (size_[2] - 1) * jump_[2]; }
    // one rank holds the whole lattice: nothing is skipped
    long sitesSkip() const { return 0; }
    long sitesSkip2d() const { return 0; }
    long *coordSkip() { return coordSkip_; }
    int getRank(int *) { return 0; }
    int getRankDim0(int) { return 0; }
    int getRankDim1(int) { return 0; }
    bool cubic() const { return size_[0] == size_[1] && size_[1] == size_[2]; }

  private:
    int dim_, halo_;
    int size_[3];
    long jump_[3];
    long coordSkip_[2];
};

// ------------------------------------------------------------------ Site / rKSite / cKSite
// A cursor holding a storage index of a lattice (ghosts included); coordinates follow from
// the index (LATfield2_Site.hpp:128-149).  first/test/next walk the interior in storage
// order (x fastest, :55-76); haloFirst/haloTest/haloNext walk the ghost sites (:99-116).
class Site {
  public:
    Site() : lat_(nullptr), index_(0) {}
    explicit Site(Lattice &lat) : lat_(&lat), index_(0) {}
    Site(Lattice &lat, long index) : lat_(&lat), index_(index) {}
    void initialize(Lattice &lat) { lat_ = &lat; index_ = 0; }
    void initialize(Lattice &lat, long index) { lat_ = &lat; index_ = index; }

    void first() { index_ = lat_->siteFirst(); }
    bool test() const { return index_ <= lat_->siteLast(); }
    void next() {
        ++index_;
        for (int d = 0; d < 3; ++d) {
            if (coordLocal(d) != lat_->sizeLocal(d)) return;  // still inside along d
            // wrapped along d: back to 0 along d, one step along d + 1
            index_ -= (long)lat_->sizeLocal(d) * lat_->jump(d);
            if (d == 2) break;
            index_ += lat_->jump(d + 1);
        }
        index_ = lat_->siteLast() + 1;
    }
    // x restricted to [offset, offset + thickness) (LATfield2_Site.hpp:78-96)
    void nextInSlice(int offset, int thickness) {
        ++index_;
        if (coordLocal(0) != offset + thickness) return;
        index_ -= thickness;
        for (int d = 1; d < 3; ++d) {
            index_ += lat_->jump(d);
            if (coordLocal(d) != lat_->sizeLocal(d)) return;
            index_ -= (long)lat_->sizeLocal(d) * lat_->jump(d);
        }
        index_ = lat_->siteLast() + 1;
    }
    void haloFirst() {
        index_ = 0;
        if (lat_->halo() == 0) index_ = lat_->sitesLocalGross();  // no ghost site
    }
    bool haloTest() const { return index_ < lat_->sitesLocalGross(); }
    void haloNext() {
        ++index_;
        // entering the interior of a row that is interior along y and z: jump over it
        if (lat_->halo() > 0 && coordLocal(0) == 0 && coordLocal(1) >= 0 && coordLocal(1) < lat_->sizeLocal(1) &&
            coordLocal(2) >= 0 && coordLocal(2) < lat_->sizeLocal(2))
            index_ += lat_->sizeLocal(0);
    }
    void indexAdvance(long number) { index_ += number; }
    long index() const { return index_; }
    void setIndex(long i) { index_ = i; }
    int coord(int dir) const { return coordLocal(dir); }
    int coordLocal(int dir) const {
        const long j = lat_->jump(dir);
        const long within = dir == 2 ? index_ : index_ % lat_->jump(dir + 1);
        return (int)(within / j) - lat_->halo();
    }
    bool setCoord(int x, int y = 0, int z = 0) {
        if (x < 0 || y < 0 || z < 0 || x >= lat_->size(0) || y >= lat_->size(1) || z >= lat_->size(2)) return false;
        index_ = (x + lat_->halo()) * lat_->jump(0) + (y + lat_->halo()) * lat_->jump(1) + (z + lat_->halo()) * lat_->jump(2);
        return true;
    }
    bool setCoord(const int *r) { return setCoord(r[0], r[1], r[2]); }
    bool setCoordLocal(const int *r) { return setCoord(r); }
    Site operator+(int dir) const { return Site(*lat_, index_ + lat_->jump(dir)); }
    Site operator-(int dir) const { return Site(*lat_, index_ - lat_->jump(dir)); }
    Site move(int dir) const { return Site(*lat_, index_ + lat_->jump(dir)); }
    Site move(int dir, int step) const { return Site(*lat_, index_ + step * lat_->jump(dir)); }
    Site move(const int *steps) const {
        long i = index_;
        for (int d = 0; d < 3; ++d) i += steps[d] * lat_->jump(d);
        return Site(*lat_, i);
    }
    // the reference's move3d steps y with jump(2) (LATfield2_Site.hpp:48-51); this one
    // steps each axis with its own jump
    Site move3d(int sx, int sy, int sz) const { return Site(*lat_, index_ + sx * lat_->jump(0) + sy * lat_->jump(1) + sz * lat_->jump(2)); }
    Lattice &lattice() const { return *lat_; }
    friend std::ostream &operator<<(std::ostream &os, const Site &x) {
        return os << " (" << x.coord(0) << " , " << x.coord(1) << " , " << x.coord(2) << ") ";
    }

  protected:
    Lattice *lat_;
    long index_;
};

// k-lattice cursor of the R2C transform: asked direction (0,1,2) = (kx, ky, kz) lives on
// storage dims (0,2,1) (LATfield2_Site.hpp:319-416)
class rKSite : public Site {
  public:
    rKSite() : Site() {}
    explicit rKSite(Lattice &lat) : Site(lat) {}
    rKSite(Lattice &lat, long index) : Site(lat, index) {}
    int coord(int dir) const { return Site::coord(map(dir)); }
    int coordLocal(int dir) const { return Site::coordLocal(map(dir)); }
    int latCoord(int dir) const { return Site::coord(dir); }
    int latCoordLocal(int dir) const { return Site::coordLocal(dir); }
    bool setCoord(int kx, int ky, int kz) {
        int r[3];
        r[map(0)] = kx;
        r[map(1)] = ky;
        r[map(2)] = kz;
        return Site::setCoord(r);
    }
    bool setCoord(const int *r) { return setCoord(r[0], r[1], r[2]); }
    rKSite operator+(int dir) const { return rKSite(*lat_, index_ + lat_->jump(map(dir))); }
    rKSite operator-(int dir) const { return rKSite(*lat_, index_ - lat_->jump(map(dir))); }

  private:
    static int map(int dir) { return dir == 0 ? 0 : (dir == 1 ? 2 : 1); }
};

// k-lattice cursor of the C2C transform: asked direction (0,1,2) = (kx, ky, kz) lives on
// storage dims (1,2,0) (LATfield2_Site.hpp:219-316)
class cKSite : public Site {
  public:
    cKSite() : Site() {}
    explicit cKSite(Lattice &lat) : Site(lat) {}
    cKSite(Lattice &lat, long index) : Site(lat, index) {}
    int coord(int dir) const { return Site::coord(map(dir)); }
    int coordLocal(int dir) const { return Site::coordLocal(map(dir)); }
    int latCoord(int dir) const { return Site::coord(dir); }
    int latCoordLocal(int dir) const { return Site::coordLocal(dir); }
    bool setCoord(int kx, int ky, int kz) {
        int r[3];
        r[map(0)] = kx;
        r[map(1)] = ky;
        r[map(2)] = kz;
        return Site::setCoord(r);
    }
    bool setCoord(const int *r) { return setCoord(r[0], r[1], r[2]); }
    cKSite operator+(int dir) const { return cKSite(*lat_, index_ + lat_->jump(map(dir))); }
    cKSite operator-(int dir) const { return cKSite(*lat_, index_ - lat_->jump(map(dir))); }

  private:
    static int map(int dir) { return dir == 0 ? 1 : (dir == 1 ? 2 : 0); }
};

// loop over the ghost sites of a lattice (looping_macro.hpp:10-21)
#define forallboundary_start(latttt, xxxx)                                                   \
    Site xxxx(latttt);                                                                       \
    for (xxxx.haloFirst(); xxxx.haloTest(); xxxx.haloNext()) {                               \
        bool inBoundary = false;                                                             \
        for (int i = 0; i < (latttt).dim(); i++)                                             \
            if (xxxx.coord(i) >= (latttt).size(i) || xxxx.coord(i) < 0) inBoundary = true;   \
        if (inBoundary == true) {
#define forallboundary_stop \
    }                       \
    }

// ------------------------------------------------------------------ Field
namespace detail {
// one device buffer and its host mirror, shared by the fields that alias it (FFT_IN_PLACE)
struct FieldStore {
    void *dev = nullptr;
    char *host = nullptr;
    int64_t bytes = 0;
    // 0: both copies equal; 1: the host copy may have been written since the last upload;
    // 2: the device copy was written since the last download (or the host copy is not
    // allocated yet)
    int state = 2;
    FieldStore() = default;
    FieldStore(const FieldStore &) = delete;
    FieldStore &operator=(const FieldStore &) = delete;
    ~FieldStore() {
        if (dev) lf2mi_device_free(dev);
        std::free(host);
    }
    void to_host() {
        if (!host) {
            host = static_cast<char *>(std::malloc((size_t)bytes));
            if (!host) lf2mi_die("LATField2d::Field: cannot allocate the host copy of the field");
        }
        lf2mi_or_die(lf2mi_copy(host, dev, bytes, LF2MI_COPY_D2H, nullptr), "Field (device to host)");
        state = 0;
    }
    void to_device() {
        lf2mi_or_die(lf2mi_copy(dev, host, bytes, LF2MI_COPY_H2D, nullptr), "Field (host to device)");
        state = 0;
    }
};
}  // namespace detail

template <class FieldType>
class Field {
  public:
    Field() : data_(nullptr), lat_(nullptr), comps_(0), rows_(0), cols_(0), nMatrix_(1), matrixSize_(0), sym_(unsymmetric) {}
    Field(Lattice &lat, int nComponents = 1) : Field() { initialize(lat, nComponents); alloc(); }
    Field(Lattice &lat, int rows, int cols, int symmetry = unsymmetric) : Field() { initialize(lat, rows, cols, symmetry); alloc(); }
    Field(Lattice &lat, int nMatrix, int rows, int cols, int symmetry) : Field() {
        initialize(lat, nMatrix, rows, cols, symmetry);
        alloc();
    }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    // LATfield2_Field_decl.hpp:620-693: a vector of `components`; a rows x cols matrix
    // (rows (rows+1)/2 components when symmetric); nMatrix such matrices
    void initialize(Lattice &lat, int nComponents = 1) { set_shape(lat, 1, nComponents, 1, unsymmetric, nComponents); }
    void initialize(Lattice &lat, int rows, int cols, int symmetry = unsymmetric) {
        set_shape(lat, 1, rows, cols, symmetry, symmetry == symmetric ? rows * (rows + 1) / 2 : rows * cols);
    }
    void initialize(Lattice &lat, int nMatrix, int rows, int cols, int symmetry) {
        set_shape(lat, nMatrix, rows, cols, symmetry, symmetry == symmetric ? rows * (rows + 1) / 2 : rows * cols);
    }
    void alloc() { alloc_elems(lat_->sitesLocalGross()); }
    // at least `size` sites (LATfield2_Field_decl.hpp:715-740)
    void alloc(long size) {
        const long n = std::max<long>(size, lat_->sitesLocalGross());
        if (st_ && st_->bytes < (int64_t)n * comps_ * (int64_t)sizeof(FieldType)) dealloc();
        alloc_elems(n);
    }
    void dealloc() {
        st_.reset();
        data_ = nullptr;
    }
    bool IsAllocated() const { return (bool)st_; }
    bool IsInitialized() const { return lat_ != nullptr; }

    // host access (LATfield2_Field_decl.hpp:760-881)
    FieldType &operator()(long index) { return hp()[index]; }
    FieldType &operator()(long index, int c) { return hp()[index * comps_ + c]; }
    FieldType &operator()(long index, int i, int j) { return hp()[index * comps_ + component(i, j)]; }
    FieldType &operator()(long index, int k, int i, int j) { return hp()[index * comps_ + component(k, i, j)]; }
    FieldType &operator()(const Site &s) { return (*this)(s.index()); }
    FieldType &operator()(const Site &s, int c) { return (*this)(s.index(), c); }
    FieldType &operator()(const Site &s, int i, int j) { return (*this)(s.index(), i, j); }
    FieldType &operator()(const Site &s, int k, int i, int j) { return (*this)(s.index(), k, i, j); }

    // component of matrix entry (i, j) (LATfield2_Field_decl.hpp:773-788): symmetric
    // matrices store the upper triangle row by row, (0,0)(0,1)(0,2)(1,1)(1,2)(2,2) for 3x3;
    // others store column-major, (i, j) at j * rows + i
    int component(int i, int j) const {
        if (sym_ == symmetric) {
            const int a = i < j ? i : j, b = i < j ? j : i;
            return b + a * rows_ - a * (a + 1) / 2;
        }
        return j * rows_ + i;
    }
    // entry (i, j) of matrix k (LATfield2_Field_decl.hpp:790-806)
    int component(int k, int i, int j) const { return component(i, j) + matrixSize_ * k; }

    // the host array (LATfield2_Field_decl.hpp:1984), brought up to date and marked as
    // possibly written
    FieldType *&data() {
        // a field that aliases another's storage may see the host copy allocated through
        // the other field, so data_ is refreshed here
        data_ = hp();
        return data_;
    }
    // mark the host copy as written (after writes through data_ kept from earlier)
    void touch_host() { hp(); }
    // force the transfers (the automatic coherence makes them optional)
    void toDevice() {
        need();
        if (st_->state == 1) st_->to_device();
    }
    void toHost() {
        need();
        if (st_->state == 2) {
            st_->to_host();
            data_ = reinterpret_cast<FieldType *>(st_->host);
        }
    }
    // the device buffer for callers that launch their own work on it: pending host writes are
    // uploaded first, and the device copy counts as written afterwards
    void *device_ptr() { return dev_write(); }
    // the same for work that only reads it
    const void *device_ptr_read() { return dev_read(); }

    // periodic ghost fill on the device
    void updateHalo() {
        lf2mi_or_die(lf2mi_ghost_fill_box(dev_write(), lat_->size(0), lat_->size(1), lat_->size(2), lat_->halo(),
                                          comps_ * (int)(sizeof(FieldType) / sizeof(Real)), lf2mi_real_code(), nullptr),
                     "Field::updateHalo");
    }

    // serial binary I/O of one process (LATfield2_Field_decl.hpp:1179-1340): the interior
    // sites in storage order (x fastest), all components of a site together
    void write(const std::string filename) {
        std::ofstream file(filename.c_str(), std::ios::binary | std::ios::trunc);
        if (!file.is_open()) {
            cout << "Latfield::Field::save - Could not open file for writing" << endl;
            cout << "Latfield::Field::save - File: " << filename << endl;
            parallel.abortRequest();
        }
        FieldType *h = hp();
        for (int z = 0; z < lat_->size(2); ++z)
            for (int y = 0; y < lat_->size(1); ++y)
                file.write(reinterpret_cast<const char *>(h + row_start(y, z) * comps_),
                           (std::streamsize)sizeof(FieldType) * comps_ * lat_->size(0));
    }
    void read(const std::string filename) {
        std::ifstream file(filename.c_str(), std::ios::binary);
        if (!file.is_open()) {
            cerr << "Latfield::Field::read - Could not open file for reading" << endl;
            cerr << "Latfield::Field::read - File: " << filename << endl;
            parallel.abortRequest();
        }
        FieldType *h = hp();
        for (int z = 0; z < lat_->size(2); ++z)
            for (int y = 0; y < lat_->size(1); ++y)
                file.read(reinterpret_cast<char *>(h + row_start(y, z) * comps_),
                          (std::streamsize)sizeof(FieldType) * comps_ * lat_->size(0));
    }

    Lattice &lattice() { return *lat_; }
    int components() const { return comps_; }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    int nMatrix() const { return nMatrix_; }
    int symmetry() const { return sym_; }

    // coherence hooks used by PlanFFT, the projections and the particles
    void *dev_read() {
        need();
        if (st_->state == 1) st_->to_device();
        return st_->dev;
    }
    void *dev_write() {
        void *d = dev_read();
        st_->state = 2;
        return d;
    }
    int64_t store_bytes() const { return st_ ? st_->bytes : 0; }
    // share another field's storage (FFT_IN_PLACE, LATfield2_PlanFFT_decl.hpp:804-824)
    template <class Other> void alias(Field<Other> &o) {
        st_ = o.store();
        data_ = reinterpret_cast<FieldType *>(st_->host);
    }
    std::shared_ptr<detail::FieldStore> store() { return st_; }

    // the host array, public as in the reference (LATfield2_Field_decl.hpp:414)
    FieldType *data_;

  private:
    void set_shape(Lattice &lat, int nMatrix, int rows, int cols, int symmetry, int matrixSize) {
        dealloc();
        lat_ = &lat;
        nMatrix_ = nMatrix;
        rows_ = rows;
        cols_ = cols;
        sym_ = symmetry;
        matrixSize_ = matrixSize;
        comps_ = matrixSize * nMatrix;
    }
    void alloc_elems(long sites) {
        if (st_) return;
        if (!lat_) lf2mi_die("LATField2d::Field::alloc: the field is not initialized");
        auto s = std::make_shared<detail::FieldStore>();
        s->bytes = (int64_t)sites * comps_ * (int64_t)sizeof(FieldType);
        lf2mi_or_die(lf2mi_device_alloc(&s->dev, s->bytes), "Field::alloc");
        lf2mi_or_die(lf2mi_zero(s->dev, s->bytes / (int64_t)sizeof(Real), lf2mi_real_code(), nullptr), "Field::alloc");
        s->state = 2;
        st_ = s;
        data_ = nullptr;
    }
    long row_start(int y, int z) const {
        const int h = lat_->halo();
        return h * lat_->jump(0) + (y + h) * lat_->jump(1) + (z + h) * lat_->jump(2);
    }
    void need() const {
        if (!st_) lf2mi_die("LATField2d::Field: the field is not allocated");
    }
    FieldType *hp() {
        detail::FieldStore *s = st_.get();
        if (!s || s->state != 1) return touch_slow();
        return reinterpret_cast<FieldType *>(s->host);
    }
    FieldType *touch_slow() {
        need();
        if (st_->state == 2) st_->to_host();
        st_->state = 1;
        data_ = reinterpret_cast<FieldType *>(st_->host);
        return data_;
    }

    Lattice *lat_;
    int comps_, rows_, cols_, nMatrix_, matrixSize_, sym_;
    std::shared_ptr<detail::FieldStore> st_;
};

// ------------------------------------------------------------------ PlanFFT
template <class compType>
class PlanFFT {
  public:
    PlanFFT() : plan_(nullptr), r_(nullptr), c_(nullptr), k_(nullptr), mem_(FFT_OUT_OF_PLACE) {}
    PlanFFT(Field<Real> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) : PlanFFT() {
        initialize(rfield, kfield, mem_type);
    }
    // C2C of two complex fields (LATfield2_PlanFFT_decl.hpp:600-719)
    PlanFFT(Field<compType> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) : PlanFFT() {
        initialize(rfield, kfield, mem_type);
    }
    ~PlanFFT() {
        if (plan_) lf2mi_fft_plan_destroy(plan_);
    }
    PlanFFT(const PlanFFT &) = delete;
    PlanFFT &operator=(const PlanFFT &) = delete;

    // checks of LATfield2_PlanFFT_decl.hpp:737-800; allocates both fields, or one shared
    // buffer for FFT_IN_PLACE (:804-824).  The transforms go through the plan's workspace, so
    // in place they read all of their input before they write any output.
    void initialize(Field<Real> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) {
        check(rfield->components(), kfield->components(), rfield->lattice());
        r_ = rfield;
        k_ = kfield;
        mem_ = mem_type;
        allocate(*rfield, *kfield);
        lf2mi_or_die(lf2mi_fft_plan_create(rfield->lattice().size(0), rfield->components(), rfield->lattice().halo(),
                                           kfield->lattice().halo(), lf2mi_real_code(), &plan_),
                     "PlanFFT::initialize");
    }
    void initialize(Field<compType> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) {
        check(rfield->components(), kfield->components(), rfield->lattice());
        c_ = rfield;
        k_ = kfield;
        mem_ = mem_type;
        allocate(*rfield, *kfield);
        lf2mi_or_die(lf2mi_fft_plan_create_c2c(rfield->lattice().size(0), rfield->components(), rfield->lattice().halo(),
                                               kfield->lattice().halo(), lf2mi_real_code(), &plan_),
                     "PlanFFT::initialize");
    }
    // FFT_FORWARD: unnormalised e^{-2 pi i k.x/N}; FFT_BACKWARD: unnormalised e^{+}, divide
    // by N^3 yourself (examples_fft.dox:123); other values do nothing, as in the reference.
    // C2C: BACKWARD is the inverse of FORWARD (times N^3) on every input.
    void execute(int fft_type) {
        if (fft_type != FFT_FORWARD && fft_type != FFT_BACKWARD) return;
        void *r, *k;
        if (fft_type == FFT_FORWARD) {
            r = r_ ? r_->dev_read() : c_->dev_read();
            k = k_->dev_write();
        } else {
            k = k_->dev_read();
            r = r_ ? r_->dev_write() : c_->dev_write();
        }
        lf2mi_or_die(lf2mi_fft_execute(plan_, fft_type, r, k, nullptr), "PlanFFT::execute");
    }
    lf2mi_fft_plan handle() const { return plan_; }

  private:
    void check(int rc, int kc, Lattice &lat) {
        if (rc != kc)
            lf2mi_die("Latfield2d::PlanFFT::initialize : fft curently work only for fields with same number of components");
        if (lat.dim() != 3 || !lat.cubic()) lf2mi_die("Latfield2d::PlanFFT::initialize : fft curently work only for 3d cubic lattice");
    }
    template <class A, class B> void allocate(Field<A> &a, Field<B> &b) {
        if (mem_ != FFT_IN_PLACE) {
            a.alloc();
            b.alloc();
            return;
        }
        const int64_t na = (int64_t)a.lattice().sitesLocalGross() * a.components() * (int64_t)sizeof(A);
        const int64_t nb = (int64_t)b.lattice().sitesLocalGross() * b.components() * (int64_t)sizeof(B);
        if (na >= nb) {
            if (a.store_bytes() < nb) a.dealloc();
            a.alloc();
            b.alias(a);
        } else {
            if (b.store_bytes() < na) b.dealloc();
            b.alloc();
            a.alias(b);
        }
    }
    lf2mi_fft_plan plan_;
    Field<Real> *r_;
    Field<compType> *c_;
    Field<compType> *k_;
    int mem_;
};

// ------------------------------------------------------------------ particles
// part_simple and part_simple_info of particles/LATfield2_particle_simple.hpp:40-67
struct part_simple {
    long ID;
    Real pos[3];
    Real vel[3];
};
struct part_simple_info {
    double mass;
    int relativistic;
    int type_name_size;
    char type_name[64];
};
struct part_simple_dataType {};

// particles/LATfield2_Particles.hpp:116-123
template <typename pInfo>
void set_parts_typename(pInfo *info, const std::string &type_name) {
    info->type_name_size = (int)type_name.size();
    std::strncpy(info->type_name, type_name.c_str(), sizeof(info->type_name) - 1);
    info->type_name[sizeof(info->type_name) - 1] = 0;
}

// The functors with a device implementation.  Particles::updateVel and moveParticles
// recognise them by address and run their kernels; called directly on the host they stop
// the program, since the particles live on the device.
template <class part, class part_info>
Real updateVel_simple(double, double, part *, double *, part_info, Field<Real> **, Site *, int, double *, double *, int) {
    lf2mi_die("updateVel_simple runs on the device: pass it to Particles::updateVel");
    return 0;
}
template <class part, class part_info>
Real updateVel_cicKick(double, double, part *, double *, part_info, Field<Real> **, Site *, int, double *, double *, int) {
    lf2mi_die("updateVel_cicKick runs on the device: pass it to Particles::updateVel");
    return 0;
}
template <class part, class part_info>
void move_particles_simple(double, double, part *, double *, part_info, Field<Real> **, Site *, int, double *, double *, int) {
    lf2mi_die("move_particles_simple runs on the device: pass it to Particles::moveParticles");
}

template <typename T, typename = void> struct has_mass_member : std::false_type {};
template <typename T> struct has_mass_member<T, decltype((void)std::declval<T>().mass, void())> : std::true_type {};

template <typename part, typename part_info, typename part_dataType>
class Particles {
  public:
    typedef Real (*updateVel_funct_t)(double, double, part *, double *, part_info, Field<Real> **, Site *, int, double *,
                                      double *, int);
    typedef void (*move_funct_t)(double, double, part *, double *, part_info, Field<Real> **, Site *, int, double *,
                                 double *, int);

    Particles() : h_(nullptr), N_(0), dirty_(false), lat_part_(nullptr) {}
    ~Particles() {
        if (h_) lf2mi_parts_destroy(h_);
    }
    Particles(const Particles &) = delete;
    Particles &operator=(const Particles &) = delete;

    // Particles::initialize (particles/LATfield2_Particles.hpp:358-428); box length 1
    void initialize(part_info info, part_dataType, Lattice *lat_part, Real boxSize[3]) {
        info_ = info;
        lat_part_ = lat_part;
        N_ = lat_part->size(0);
        if (!lat_part->cubic()) lf2mi_die("LATfield2::Particles: the particle lattice must be cubic");
        for (int i = 0; i < 3; ++i)
            if (boxSize[i] != (Real)1)
                lf2mi_die("LATfield2::Particles: the MI355X path uses box length 1 (the binning of getPartCoord assumes it)");
        staged_.clear();
        dirty_ = true;
    }
    // addParticle_global (particles/LATfield2_Particles.hpp:501-519): staged on the host,
    // uploaded on the next device operation; negative cells are rejected as in the reference
    bool addParticle_global(const part &p) {
        for (int d = 0; d < 3; ++d)
            if ((int)std::floor(p.pos[d] * N_) % N_ < 0) return false;
        staged_.push_back(p);
        dirty_ = true;
        return true;
    }
    // the cell of a position (particles/LATfield2_Particles.hpp:464-467)
    void getPartCoord(Real pos[3], int coord[3]) {
        for (int d = 0; d < 3; ++d) coord[d] = (int)std::floor(pos[d] * N_) % N_;
    }
    bool getPartCoordLocal(Real pos[3], int coord[3]) {
        getPartCoord(pos, coord);
        return true;
    }
    long numParticles() {
        sync();
        int64_t n = 0;
        lf2mi_or_die(lf2mi_parts_count(h_, &n), "Particles::numParticles");
        return (long)n;
    }
    part_info &parts_info() { return info_; }
    double res() const { return 1.0 / N_; }
    Lattice &lattice() { return *lat_part_; }

    // updateVel (particles/LATfield2_Particles.hpp:753-893) with a device functor: returns
    // sqrt(max |v|^2) after the update; output[i] receives slot i of the functor's outputs
    // reduced by reduce_type[i]
    Real updateVel(updateVel_funct_t funct, double dtau, Field<Real> **fields = NULL, int nfields = 0, double *params = NULL,
                   double *output = NULL, int *reduce_type = NULL, int noutput = 0) {
        (void)params;
        sync();
        if (funct == &updateVel_simple<part, part_info>) {
            // the identity on the velocities
        } else if (funct == &updateVel_cicKick<part, part_info>) {
            if (nfields < 1 || fields == NULL || fields[0]->components() != 1)
                lf2mi_die("LATfield2::Particles::updateVel: updateVel_cicKick needs the potential as fields[0]");
            lf2mi_or_die(lf2mi_gather_kick_drift(h_, fields[0]->dev_read(), fields[0]->lattice().halo(), dtau, 0.0, nullptr,
                                                 nullptr),
                         "Particles::updateVel");
        } else {
            lf2mi_die("LATfield2::Particles::updateVel: this functor has no device implementation (updateVel_simple and "
                      "updateVel_cicKick have)");
        }
        return (Real)std::sqrt(reduce(output, reduce_type, noutput));
    }
    // moveParticles (particles/LATfield2_Particles.hpp:1811-2803) with a device functor
    void moveParticles(move_funct_t funct, double dtau, Field<Real> **fields = NULL, int nfields = 0, double *params = NULL,
                       double *output = NULL, int *reduce_type = NULL, int noutput = 0) {
        (void)fields;
        (void)nfields;
        (void)params;
        sync();
        if (funct != &move_particles_simple<part, part_info>)
            lf2mi_die("LATfield2::Particles::moveParticles: this functor has no device implementation "
                      "(move_particles_simple has)");
        lf2mi_or_die(lf2mi_gather_kick_drift(h_, nullptr, 1, 0.0, dtau, nullptr, nullptr), "Particles::moveParticles");
        if (noutput > 0) reduce(output, reduce_type, noutput);
    }
    // copy all particles back (container order): the whole structs as added, with pos,
    // vel and ID brought up to date
    void getParticles(std::vector<part> &out) {
        sync();
        const long n = numParticles();
        out.resize((size_t)n);
        int64_t got = 0;
        lf2mi_or_die(lf2mi_parts_download_aos(h_, out.data(), (int64_t)n, (int64_t)sizeof(part), lf2mi_real_code(),
                                              (int64_t)offsetof(part, pos), (int64_t)offsetof(part, vel),
                                              (int64_t)offsetof(part, ID), &got, nullptr),
                     "Particles::getParticles");
        out.resize((size_t)got);
    }
    lf2mi_parts handle() {
        sync();
        return h_;
    }
    double mass() const { return info_.mass; }
    // the mass a projection uses (projections.hpp:110-126): without oset the one detected at
    // initialize (a `mass` member of part, else part_info's mass, Particles.hpp:398-427);
    // with oset the double at that byte offset of part_info (FROM_INFO) or of every
    // particle (FROM_PART)
    void select_mass(size_t *oset, int flag_where) {
        sync();
        if (oset == nullptr) {
            if (has_mass_member<part>::value)
                lf2mi_or_die(lf2mi_parts_mass_from(h_, mass_source<part>(), mass_offset<part>(), nullptr), "projection mass");
            else
                use_global_mass(info_.mass);
        } else if (flag_where == FROM_INFO) {
            double m;
            std::memcpy(&m, reinterpret_cast<const char *>(&info_) + *oset, sizeof(double));
            use_global_mass(m);
        } else {
            lf2mi_or_die(lf2mi_parts_mass_from(h_, LF2MI_MASS_PART, (int64_t)*oset, nullptr), "projection mass");
        }
    }

    // the vector a vector or tensor projection uses (projections.hpp:453-459): without oset
    // the velocity, with oset the Real[3] at byte offset oset[1] of every particle
    void select_vector(size_t *oset) {
        sync();
        lf2mi_or_die(lf2mi_parts_vector_from(h_, oset ? (int64_t)oset[1] : -1, lf2mi_real_code()), "projection vector");
    }

  private:
    // the outputs of the device functors (slot i = quantity i) and max |v|^2
    double reduce(double *output, int *reduce_type, int noutput) {
        static const int kSlots = 6;
        if (noutput > kSlots || (noutput > 0 && (!output || !reduce_type)))
            lf2mi_die("LATfield2::Particles: the device functors have 6 output slots (|v|^2, m v_x, m v_y, m v_z, m, "
                      "m |v|^2 / 2)");
        int q[kSlots + 1], rt[kSlots + 1];
        double out[kSlots + 1];
        q[0] = LF2MI_PQ_V2;
        rt[0] = MAX;
        out[0] = 0.0;
        for (int i = 0; i < noutput; ++i) {
            q[i + 1] = i;
            rt[i + 1] = reduce_type[i];
            out[i + 1] = output[i];
        }
        lf2mi_or_die(lf2mi_parts_reduce(h_, noutput + 1, q, rt, out, nullptr), "Particles (output reduction)");
        for (int i = 0; i < noutput; ++i) output[i] = out[i + 1];
        return numParticles() > 0 ? out[0] : 0.0;
    }
    // the staged structs go to the device whole (a payload of sizeof(part) bytes moves with
    // each particle), read by their members' offsets
    void sync() {
        if (!dirty_) return;
        if (h_) lf2mi_parts_destroy(h_);
        h_ = nullptr;
        static_assert(sizeof(part) % 8 == 0, "part structs are kept whole in 8-byte words");
        static_assert(sizeof(((part *)nullptr)->ID) == 8, "the part ID is a 64-bit long");
        const bool pm = has_mass_member<part>::value;
        lf2mi_or_die(lf2mi_parts_create(N_, (int64_t)staged_.size(), lf2mi_real_code(), LF2MI_PARTS_ID | (pm ? LF2MI_PARTS_MASS : 0), &h_),
                     "Particles::initialize");
        lf2mi_or_die(lf2mi_parts_set_mass(h_, info_.mass), "Particles::initialize");
        lf2mi_or_die(lf2mi_parts_set_payload(h_, (int)sizeof(part)), "Particles::initialize");
        lf2mi_or_die(lf2mi_parts_upload_aos(h_, (int64_t)staged_.size(), staged_.data(), (int64_t)sizeof(part),
                                            lf2mi_real_code(), (int64_t)offsetof(part, pos), (int64_t)offsetof(part, vel),
                                            (int64_t)offsetof(part, ID), pm && mass_source<part>() == LF2MI_MASS_PART
                                                                             ? mass_offset<part>() : -1,
                                            nullptr, nullptr),
                     "Particles::addParticle_global");
        staged_.clear();
        dirty_ = false;
    }
    template <typename P>
    static int64_t mass_offset() {
        if constexpr (has_mass_member<P>::value) {
            static_assert(sizeof(((P *)nullptr)->mass) == sizeof(double) || sizeof(((P *)nullptr)->mass) == sizeof(float),
                          "a per-particle mass is a double or a float");
            return (int64_t)offsetof(P, mass);
        } else {
            return -1;
        }
    }
    // a mass member is read with its declared type (a float one: Real mass under SINGLE)
    template <typename P>
    static int mass_source() {
        if constexpr (has_mass_member<P>::value) {
            return sizeof(((P *)nullptr)->mass) == sizeof(float) ? LF2MI_MASS_PART_F32 : LF2MI_MASS_PART;
        } else {
            return LF2MI_MASS_INFO;
        }
    }
    void use_global_mass(double m) {
        lf2mi_or_die(lf2mi_parts_set_mass(h_, m), "projection mass");
        lf2mi_or_die(lf2mi_parts_mass_from(h_, LF2MI_MASS_INFO, -1, nullptr), "projection mass");
    }

    lf2mi_parts h_;
    int N_;
    bool dirty_;
    Lattice *lat_part_;
    part_info info_;
    std::vector<part> staged_;
};

// ------------------------------------------------------------------ projections
inline void projection_init(Field<Real> *f) {
    lf2mi_or_die(lf2mi_zero(f->dev_write(), f->lattice().sitesLocalGross() * f->components(), lf2mi_real_code(), nullptr),
                 "projection_init");
}

namespace detail {
template <typename P>
void project(P *parts, Field<Real> *f, int kind, const char *where) {
    lf2mi_or_die(lf2mi_deposit(parts->handle(), kind, f->dev_write(), f->lattice().halo(), LF2MI_DEP_UNFOLDED, LF2MI_ACC_NATIVE,
                               nullptr),
                 where);
}
inline void comm(Field<Real> *f, int kind, const char *where) {
    lf2mi_or_die(lf2mi_projection_comm(f->dev_write(), f->lattice().size(0), f->lattice().halo(), kind, lf2mi_real_code(), nullptr),
                 where);
}
}  // namespace detail

template <typename part, typename part_info, typename part_dataType>
void scalarProjectionCIC_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *rho, size_t *oset = NULL,
                                 int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    detail::project(parts, rho, LF2MI_PROJ_SCALAR, "scalarProjectionCIC_proj");
}
inline void scalarProjectionCIC_comm(Field<Real> *rho) { detail::comm(rho, LF2MI_PROJ_SCALAR, "scalarProjectionCIC_comm"); }

template <typename part, typename part_info, typename part_dataType>
void vectorProjectionCIC_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *B, size_t *oset = NULL,
                                 int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    detail::project(parts, B, LF2MI_PROJ_VECTOR, "vectorProjectionCIC_project");
}
inline void vectorProjectionCIC_comm(Field<Real> *B) { detail::comm(B, LF2MI_PROJ_VECTOR, "vectorProjectionCIC_comm"); }

template <typename part, typename part_info, typename part_dataType>
void VecVecProjectionCIC_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *T, size_t *oset = NULL,
                                 int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    detail::project(parts, T, LF2MI_PROJ_VECVEC, "VecVecProjectionCIC_project");
}
inline void VecVecProjectionCIC_comm(Field<Real> *T) { detail::comm(T, LF2MI_PROJ_VECVEC, "VecVecProjectionCIC_comm"); }

// the CIC-NGP hybrids (projections.hpp:405-976)
template <typename part, typename part_info, typename part_dataType>
void vectorProjectionCICNGP_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *vel,
                                    size_t *oset = NULL, int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    detail::project(parts, vel, LF2MI_PROJ_VECTOR_NGP, "vectorProjectionCICNGP_project");
}
inline void vectorProjectionCICNGP_comm(Field<Real> *vel) {
    detail::comm(vel, LF2MI_PROJ_VECTOR_NGP, "vectorProjectionCICNGP_comm");
}
template <typename part, typename part_info, typename part_dataType>
void symtensorProjectionCICNGP_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *Tij,
                                       size_t *oset = NULL, int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    detail::project(parts, Tij, LF2MI_PROJ_SYMTENSOR_NGP, "symtensorProjectionCICNGP_project");
}
inline void symtensorProjectionCICNGP_comm(Field<Real> *Tij) {
    detail::comm(Tij, LF2MI_PROJ_SYMTENSOR_NGP, "symtensorProjectionCICNGP_comm");
}

}  // namespace LATfield2

#endif
