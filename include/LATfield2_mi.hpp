/*
 * LATfield2_mi.hpp -- C++ mirror of the LATfield2 API for the particle-mesh path,
 * running on one MI355X through the C ABI of lf2mi.h.
 *
 * Same names and argument meaning as the reference (namespace LATfield2):
 *   Lattice                      LATfield2_Lattice_decl.hpp:13-218
 *   Site, rKSite, cKSite         LATfield2_Site_decl.hpp:13-277 (rKSite dims {0,2,1},
 *                                cKSite dims {1,2,0})
 *   Imag                         Imag.hpp:24-110
 *   Field<Real>, Field<Imag>     LATfield2_Field_decl.hpp:61-430 (AoS, symmetric 3x3 -> 6)
 *   Field::updateHalo            LATfield2_Field_decl.hpp:889-935
 *   PlanFFT<Imag>, execute       LATfield2_PlanFFT_decl.hpp:83-1203 (R2C and C2C; the
 *                                C2C backward is the inverse of the forward, which the
 *                                reference's single-rank backward is not)
 *   Particles<part,info,dtype>   particles/LATfield2_Particles.hpp:131-344
 *   projection_init, scalar/vector/VecVec ProjectionCIC _project/_comm
 *                                particles/projections.hpp:31-1814
 * Differences a user of the reference must know:
 *   - field data live in device memory (data() is a device pointer).  operator()
 *     reads and writes a host mirror; call toDevice() after host writes and toHost()
 *     before host reads, around Site loops.
 *   - updateVel/moveParticles take the device functors gpu_gather_kick / gpu_drift
 *     (CIC gather of -grad phi + kick; pos += dtau vel) instead of host function
 *     pointers, which cannot run on the GPU.
 *   - errors print the library's message and exit(-1), as the reference aborts.
 * Precision: Real is double, or float when SINGLE is defined, as in Imag.hpp:9-15.
 * Link with -llf2mi (latfield2_amd/liblf2mi.so).
 */
#ifndef LATFIELD2_MI_HPP
#define LATFIELD2_MI_HPP

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "lf2mi.h"

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

// ------------------------------------------------------------------ Imag
class Imag {
  public:
    Imag() : re_(0), im_(0) {}
    Imag(Real re, Real im = 0) : re_(re), im_(im) {}
    Real &real() { return re_; }
    Real &imag() { return im_; }
    Real real() const { return re_; }
    Real imag() const { return im_; }
    Imag conj() const { return Imag(re_, -im_); }
    Real norm() const { return re_ * re_ + im_ * im_; }
    Imag operator+(const Imag &o) const { return Imag(re_ + o.re_, im_ + o.im_); }
    Imag operator-(const Imag &o) const { return Imag(re_ - o.re_, im_ - o.im_); }
    Imag operator*(const Imag &o) const { return Imag(re_ * o.re_ - im_ * o.im_, re_ * o.im_ + im_ * o.re_); }
    Imag operator*(Real s) const { return Imag(re_ * s, im_ * s); }
    Imag operator/(Real s) const { return Imag(re_ / s, im_ / s); }
    Imag operator/(const Imag &o) const {
        const Real d = o.norm();
        return Imag((re_ * o.re_ + im_ * o.im_) / d, (im_ * o.re_ - re_ * o.im_) / d);
    }
    Imag &operator+=(const Imag &o) { re_ += o.re_; im_ += o.im_; return *this; }
    Imag &operator*=(Real s) { re_ *= s; im_ *= s; return *this; }
    Imag &operator=(Real s) { re_ = s; im_ = 0; return *this; }

  private:
    Real re_, im_;
};

// ------------------------------------------------------------------ Lattice
class Lattice {
  public:
    Lattice() : dim_(0), halo_(0) {
        for (int i = 0; i < 3; ++i) { size_[i] = 0; jump_[i] = 0; }
    }
    Lattice(int dim, const int size, int halo) : Lattice() { initialize(dim, size, halo); }
    Lattice(int dim, const int *size, int halo) : Lattice() { initialize(dim, size, halo); }

    void initialize(int dim, const int size, int halo) {
        int s[3] = {size, size, size};
        initialize(dim, s, halo);
    }
    void initialize(int dim, const int *size, int halo) {
        if (dim != 3) {
            std::fprintf(stderr, "LATfield2::Lattice: the MI355X PM path supports 3 dimensions\n");
            std::exit(-1);
        }
        dim_ = dim;
        halo_ = halo;
        for (int i = 0; i < 3; ++i) size_[i] = size[i];
        jump_[0] = 1;
        for (int i = 1; i < 3; ++i) jump_[i] = jump_[i - 1] * (size_[i - 1] + 2 * halo_);
    }
    // (N/2+1, N, N) k-lattice of the R2C transform (LATfield2_Lattice.hpp:225-259)
    void initializeRealFFT(Lattice &lat_real, int halo) {
        if (lat_real.dim() != 3 || lat_real.size(0) != lat_real.size(1) || lat_real.size(1) != lat_real.size(2)) {
            std::fprintf(stderr, "Latfield2d::Lattice::initializeRealFFT : fft curently work only for 3d cubic lattice\n");
            std::exit(-1);
        }
        int s[3] = {lat_real.size(0) / 2 + 1, lat_real.size(0), lat_real.size(0)};
        initialize(3, s, halo);
    }
    // N^3 k-lattice of the C2C transform (LATfield2_Lattice.hpp:260-297)
    void initializeComplexFFT(Lattice &lat_real, int halo) {
        if (lat_real.dim() != 3 || lat_real.size(0) != lat_real.size(1) || lat_real.size(1) != lat_real.size(2)) {
            std::fprintf(stderr, "Latfield2d::Lattice::initializeRealFFT : fft curently work only for 3d cubic lattice\n");
            std::exit(-1);
        }
        initialize(3, lat_real.size(0), halo);
    }

    int dim() const { return dim_; }
    int size(int i) const { return size_[i]; }
    const int *size() const { return size_; }
    int sizeLocal(int i) const { return size_[i]; }
    int halo() const { return halo_; }
    long jump(int i) const { return jump_[i]; }
    long sites() const { return (long)size_[0] * size_[1] * size_[2]; }
    long sitesLocal() const { return sites(); }
    long sitesGross() const { return (long)(size_[0] + 2 * halo_) * (size_[1] + 2 * halo_) * (size_[2] + 2 * halo_); }
    long sitesLocalGross() const { return sitesGross(); }
    long siteFirst() const { return halo_ * (jump_[0] + jump_[1] + jump_[2]); }
    long siteLast() const { return sitesGross() - 1 - siteFirst(); }

  private:
    int dim_, halo_;
    int size_[3];
    long jump_[3];
};

// ------------------------------------------------------------------ Site / rKSite
// Cursor over the interior in storage order (x fastest), as Site::first/test/next
// (LATfield2_Site.hpp:55-76); operator+/- step along a lattice direction.
class Site {
  public:
    Site() : lat_(nullptr), index_(0) { c_[0] = c_[1] = c_[2] = 0; }
    explicit Site(Lattice &lat) : Site() { initialize(lat); }
    void initialize(Lattice &lat) { lat_ = &lat; first(); }
    void first() {
        c_[0] = c_[1] = c_[2] = 0;
        index_ = lat_->siteFirst();
    }
    bool test() const { return c_[2] < lat_->size(2); }
    void next() {
        ++c_[0];
        ++index_;
        if (c_[0] == lat_->size(0)) {
            c_[0] = 0;
            ++c_[1];
            if (c_[1] == lat_->size(1)) {
                c_[1] = 0;
                ++c_[2];
            }
            index_ = storage(c_[0], c_[1], c_[2]);
        }
    }
    long index() const { return index_; }
    int coord(int dir) const { return c_[dir]; }
    int coordLocal(int dir) const { return c_[dir]; }
    bool setCoord(int x, int y, int z) {
        if (x < 0 || y < 0 || z < 0 || x >= lat_->size(0) || y >= lat_->size(1) || z >= lat_->size(2)) return false;
        c_[0] = x;
        c_[1] = y;
        c_[2] = z;
        index_ = storage(x, y, z);
        return true;
    }
    bool setCoord(const int *r) { return setCoord(r[0], r[1], r[2]); }
    Site operator+(int dir) const { return shifted(dir, +1); }
    Site operator-(int dir) const { return shifted(dir, -1); }
    Lattice &lattice() const { return *lat_; }

  protected:
    long storage(int x, int y, int z) const {
        const int h = lat_->halo();
        return (x + h) * lat_->jump(0) + (y + h) * lat_->jump(1) + (z + h) * lat_->jump(2);
    }
    Site shifted(int dir, int s) const {
        Site o(*this);
        o.c_[dir] += s;
        o.index_ += s * lat_->jump(dir);
        return o;
    }
    Lattice *lat_;
    long index_;
    int c_[3];
};

// k-lattice cursor: asked direction (0,1,2) = (kx, ky, kz) maps to storage dims (0,2,1)
// (LATfield2_Site.hpp:319-363)
class rKSite : public Site {
  public:
    rKSite() : Site() {}
    explicit rKSite(Lattice &lat) : Site(lat) {}
    int coord(int dir) const { return c_[map(dir)]; }
    rKSite operator+(int dir) const { return rks(Site::operator+(map(dir))); }
    rKSite operator-(int dir) const { return rks(Site::operator-(map(dir))); }

  private:
    static int map(int dir) { return dir == 0 ? 0 : (dir == 1 ? 2 : 1); }
    static rKSite rks(const Site &s) {
        rKSite r;
        static_cast<Site &>(r) = s;
        return r;
    }
};

// C2C k-lattice cursor: asked direction (0,1,2) = (kx, ky, kz) maps to storage dims (1,2,0)
// (LATfield2_Site.hpp:219-263)
class cKSite : public Site {
  public:
    cKSite() : Site() {}
    explicit cKSite(Lattice &lat) : Site(lat) {}
    int coord(int dir) const { return c_[map(dir)]; }
    int coordLocal(int dir) const { return c_[map(dir)]; }
    cKSite operator+(int dir) const { return cks(Site::operator+(map(dir))); }
    cKSite operator-(int dir) const { return cks(Site::operator-(map(dir))); }

  private:
    static int map(int dir) { return dir == 0 ? 1 : (dir == 1 ? 2 : 0); }
    static cKSite cks(const Site &s) {
        cKSite r;
        static_cast<Site &>(r) = s;
        return r;
    }
};

// ------------------------------------------------------------------ Field
template <class FieldType>
class Field {
  public:
    Field() : lat_(nullptr), comps_(0), rows_(0), cols_(0), sym_(unsymmetric), dev_(nullptr) {}
    Field(Lattice &lat, int nComponents = 1) : Field() { initialize(lat, nComponents); alloc(); }
    Field(Lattice &lat, int rows, int cols, int symmetry) : Field() { initialize(lat, rows, cols, symmetry); alloc(); }
    ~Field() { dealloc(); }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    void initialize(Lattice &lat, int nComponents = 1) {
        lat_ = &lat;
        comps_ = nComponents;
        rows_ = nComponents;
        cols_ = 1;
        sym_ = unsymmetric;
    }
    // 3x3 symmetric -> 6 components (LATfield2_Field_decl.hpp:773-788)
    void initialize(Lattice &lat, int rows, int cols, int symmetry) {
        lat_ = &lat;
        rows_ = rows;
        cols_ = cols;
        sym_ = symmetry;
        comps_ = (symmetry == symmetric && rows == cols) ? rows * (rows + 1) / 2 : rows * cols;
    }
    void alloc() {
        if (dev_) return;
        const int64_t n = (int64_t)lat_->sitesLocalGross() * comps_;
        lf2mi_or_die(lf2mi_device_alloc(&dev_, n * (int64_t)sizeof(FieldType)), "Field::alloc");
        lf2mi_or_die(lf2mi_zero(dev_, n * (int64_t)(sizeof(FieldType) / sizeof(Real)), lf2mi_real_code(), nullptr),
                     "Field::alloc");
        host_.assign((size_t)n, FieldType());
    }
    void dealloc() {
        if (dev_) lf2mi_device_free(dev_);
        dev_ = nullptr;
        host_.clear();
    }

    // host mirror access
    FieldType &operator()(long index) { return host_[(size_t)index]; }
    FieldType &operator()(long index, int c) { return host_[(size_t)(index * comps_ + c)]; }
    FieldType &operator()(const Site &s) { return host_[(size_t)(s.index() * comps_)]; }
    FieldType &operator()(const Site &s, int c) { return host_[(size_t)(s.index() * comps_ + c)]; }
    FieldType &operator()(const Site &s, int i, int j) { return (*this)(s, component(i, j)); }
    int component(int i, int j) const {
        if (sym_ == symmetric) {
            if (i > j) { const int t = i; i = j; j = t; }
            return i * rows_ - i * (i - 1) / 2 + (j - i);
        }
        return i * cols_ + j;
    }

    void toDevice() {
        lf2mi_or_die(lf2mi_copy(dev_, host_.data(), (int64_t)(host_.size() * sizeof(FieldType)), LF2MI_COPY_H2D, nullptr),
                     "Field::toDevice");
    }
    void toHost() {
        lf2mi_or_die(lf2mi_copy(host_.data(), dev_, (int64_t)(host_.size() * sizeof(FieldType)), LF2MI_COPY_D2H, nullptr),
                     "Field::toHost");
    }

    // periodic ghost fill on the device
    void updateHalo() {
        lf2mi_or_die(lf2mi_ghost_fill(dev_, lat_->size(0), lat_->halo(),
                                      comps_ * (int)(sizeof(FieldType) / sizeof(Real)), lf2mi_real_code(), nullptr),
                     "Field::updateHalo");
    }

    FieldType *data() { return static_cast<FieldType *>(dev_); }
    void *&device_ptr() { return dev_; }
    FieldType *host() { return host_.data(); }
    Lattice &lattice() { return *lat_; }
    int components() const { return comps_; }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    int symmetry() const { return sym_; }

  private:
    Lattice *lat_;
    int comps_, rows_, cols_, sym_;
    void *dev_;
    std::vector<FieldType> host_;
};

// ------------------------------------------------------------------ PlanFFT
template <class compType>
class PlanFFT {
  public:
    PlanFFT() : plan_(nullptr), r_(nullptr), c_(nullptr), k_(nullptr), stage_(nullptr), mem_(FFT_OUT_OF_PLACE) {}
    PlanFFT(Field<Real> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) : PlanFFT() {
        initialize(rfield, kfield, mem_type);
    }
    // C2C of two complex fields (LATfield2_PlanFFT_decl.hpp:600-719)
    PlanFFT(Field<compType> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) : PlanFFT() {
        initialize(rfield, kfield, mem_type);
    }
    ~PlanFFT() {
        if (plan_) lf2mi_fft_plan_destroy(plan_);
        if (stage_) lf2mi_device_free(stage_);
    }
    PlanFFT(const PlanFFT &) = delete;
    PlanFFT &operator=(const PlanFFT &) = delete;

    // checks as LATfield2_PlanFFT_decl.hpp:737-800, allocates both fields (:804-824)
    void initialize(Field<Real> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) {
        if (rfield->components() != kfield->components()) {
            std::fprintf(stderr, "Latfield2d::PlanFFT::initialize : fft curently work only for fields with same number of components\n");
            std::exit(-1);
        }
        r_ = rfield;
        k_ = kfield;
        mem_ = mem_type;
        rfield->alloc();
        kfield->alloc();
        lf2mi_or_die(lf2mi_fft_plan_create(rfield->lattice().size(0), rfield->components(), rfield->lattice().halo(),
                                           kfield->lattice().halo(), lf2mi_real_code(), &plan_),
                     "PlanFFT::initialize");
    }
    void initialize(Field<compType> *rfield, Field<compType> *kfield, const int mem_type = FFT_OUT_OF_PLACE) {
        if (rfield->components() != kfield->components()) {
            std::fprintf(stderr, "Latfield2d::PlanFFT::initialize : fft curently work only for fields with same number of components\n");
            std::exit(-1);
        }
        c_ = rfield;
        k_ = kfield;
        mem_ = mem_type;
        rfield->alloc();
        kfield->alloc();
        lf2mi_or_die(lf2mi_fft_plan_create_c2c(rfield->lattice().size(0), rfield->components(), rfield->lattice().halo(),
                                               kfield->lattice().halo(), lf2mi_real_code(), &plan_),
                     "PlanFFT::initialize");
    }
    // FFT_FORWARD: unnormalised e^{-2 pi i k.x/N}; FFT_BACKWARD: unnormalised e^{+}, divide
    // by N^3 yourself (examples_fft.dox:123); other values do nothing, as in the reference.
    // C2C: BACKWARD is the inverse of FORWARD (times N^3) on every input.
    void execute(int fft_type) {
        void *r = r_ ? r_->device_ptr() : c_->device_ptr();
        lf2mi_or_die(lf2mi_fft_execute(plan_, fft_type, r, k_->device_ptr(), nullptr), "PlanFFT::execute");
    }

  private:
    lf2mi_fft_plan plan_;
    Field<Real> *r_;
    Field<compType> *c_;
    Field<compType> *k_;
    void *stage_;
    int mem_;
};

// ------------------------------------------------------------------ particles
// part_simple and part_simple_info of particles/LATfield2_particle_simple.hpp:40-45
struct part_simple {
    long ID;
    Real pos[3];
    Real vel[3];
};
struct part_simple_info {
    double mass;
    int relativistic;
    char type_name[64];
};
struct part_simple_dataType {};

// device functor tags for updateVel / moveParticles
struct gpu_gather_kick {};  // vel += dtau * sum_corners w (-grad phi)(corner)
struct gpu_drift {};        // pos += dtau * vel, periodic wrap, re-binning

template <typename T, typename = void> struct has_mass_member : std::false_type {};
template <typename T> struct has_mass_member<T, decltype((void)std::declval<T>().mass, void())> : std::true_type {};

template <typename part, typename part_info, typename part_dataType>
class Particles {
  public:
    Particles() : h_(nullptr), N_(0), dirty_(false) {}
    ~Particles() {
        if (h_) lf2mi_parts_destroy(h_);
    }
    Particles(const Particles &) = delete;
    Particles &operator=(const Particles &) = delete;

    // Particles::initialize (particles/LATfield2_Particles.hpp:358-428); box length 1
    void initialize(part_info info, part_dataType, Lattice *lat_part, Real boxSize[3]) {
        info_ = info;
        N_ = lat_part->size(0);
        for (int i = 0; i < 3; ++i)
            if (boxSize[i] != (Real)1) {
                std::fprintf(stderr, "LATfield2::Particles: the MI355X path uses box length 1 (the binning of "
                                     "getPartCoord assumes it)\n");
                std::exit(-1);
            }
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
    long numParticles() {
        sync();
        int64_t n = 0;
        lf2mi_or_die(lf2mi_parts_count(h_, &n), "Particles::numParticles");
        return (long)n;
    }
    Real updateVel(gpu_gather_kick, double dtau, Field<Real> **fields, int nfields = 1) {
        sync();
        if (nfields < 1) lf2mi_or_die(LF2MI_EINVAL, "Particles::updateVel (a potential field is required)");
        double vmax2 = 0.0;
        lf2mi_or_die(lf2mi_gather_kick_drift(h_, fields[0]->device_ptr(), fields[0]->lattice().halo(), dtau, 0.0,
                                             &vmax2, nullptr),
                     "Particles::updateVel");
        return (Real)std::sqrt(vmax2);
    }
    void moveParticles(gpu_drift, double dtau) {
        sync();
        lf2mi_or_die(lf2mi_gather_kick_drift(h_, nullptr, 1, 0.0, dtau, nullptr, nullptr), "Particles::moveParticles");
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
                lf2mi_or_die(lf2mi_parts_mass_from(h_, mass_source<part>(), mass_offset<part>(), nullptr),
                             "projection mass");
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
    part_info info_;
    std::vector<part> staged_;
};

// ------------------------------------------------------------------ projections
inline void projection_init(Field<Real> *f) {
    lf2mi_or_die(lf2mi_zero(f->device_ptr(), f->lattice().sitesLocalGross() * f->components(), lf2mi_real_code(), nullptr),
                 "projection_init");
}

template <typename part, typename part_info, typename part_dataType>
void scalarProjectionCIC_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *rho, size_t *oset = NULL,
                                 int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    lf2mi_or_die(lf2mi_deposit(parts->handle(), LF2MI_PROJ_SCALAR, rho->device_ptr(), rho->lattice().halo(),
                               LF2MI_DEP_UNFOLDED, LF2MI_ACC_NATIVE, nullptr),
                 "scalarProjectionCIC_proj");
}
inline void scalarProjectionCIC_comm(Field<Real> *rho) {
    lf2mi_or_die(lf2mi_projection_comm(rho->device_ptr(), rho->lattice().size(0), rho->lattice().halo(),
                                       LF2MI_PROJ_SCALAR, lf2mi_real_code(), nullptr),
                 "scalarProjectionCIC_comm");
}
template <typename part, typename part_info, typename part_dataType>
void vectorProjectionCIC_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *B, size_t *oset = NULL,
                                 int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    lf2mi_or_die(lf2mi_deposit(parts->handle(), LF2MI_PROJ_VECTOR, B->device_ptr(), B->lattice().halo(),
                               LF2MI_DEP_UNFOLDED, LF2MI_ACC_NATIVE, nullptr),
                 "vectorProjectionCIC_project");
}
inline void vectorProjectionCIC_comm(Field<Real> *B) {
    lf2mi_or_die(lf2mi_projection_comm(B->device_ptr(), B->lattice().size(0), B->lattice().halo(), LF2MI_PROJ_VECTOR,
                                       lf2mi_real_code(), nullptr),
                 "vectorProjectionCIC_comm");
}
template <typename part, typename part_info, typename part_dataType>
void VecVecProjectionCIC_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *T, size_t *oset = NULL,
                                 int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    lf2mi_or_die(lf2mi_deposit(parts->handle(), LF2MI_PROJ_VECVEC, T->device_ptr(), T->lattice().halo(),
                               LF2MI_DEP_UNFOLDED, LF2MI_ACC_NATIVE, nullptr),
                 "VecVecProjectionCIC_project");
}
inline void VecVecProjectionCIC_comm(Field<Real> *T) {
    lf2mi_or_die(lf2mi_projection_comm(T->device_ptr(), T->lattice().size(0), T->lattice().halo(), LF2MI_PROJ_VECVEC,
                                       lf2mi_real_code(), nullptr),
                 "VecVecProjectionCIC_comm");
}
// the CIC-NGP hybrids (projections.hpp:405-976)
template <typename part, typename part_info, typename part_dataType>
void vectorProjectionCICNGP_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *vel,
                                    size_t *oset = NULL, int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    lf2mi_or_die(lf2mi_deposit(parts->handle(), LF2MI_PROJ_VECTOR_NGP, vel->device_ptr(), vel->lattice().halo(),
                               LF2MI_DEP_UNFOLDED, LF2MI_ACC_NATIVE, nullptr),
                 "vectorProjectionCICNGP_project");
}
inline void vectorProjectionCICNGP_comm(Field<Real> *vel) {
    lf2mi_or_die(lf2mi_projection_comm(vel->device_ptr(), vel->lattice().size(0), vel->lattice().halo(),
                                       LF2MI_PROJ_VECTOR_NGP, lf2mi_real_code(), nullptr),
                 "vectorProjectionCICNGP_comm");
}
template <typename part, typename part_info, typename part_dataType>
void symtensorProjectionCICNGP_project(Particles<part, part_info, part_dataType> *parts, Field<Real> *Tij,
                                       size_t *oset = NULL, int flag_where = FROM_INFO) {
    parts->select_mass(oset, flag_where);
    parts->select_vector(oset);
    lf2mi_or_die(lf2mi_deposit(parts->handle(), LF2MI_PROJ_SYMTENSOR_NGP, Tij->device_ptr(), Tij->lattice().halo(),
                               LF2MI_DEP_UNFOLDED, LF2MI_ACC_NATIVE, nullptr),
                 "symtensorProjectionCICNGP_project");
}
inline void symtensorProjectionCICNGP_comm(Field<Real> *Tij) {
    lf2mi_or_die(lf2mi_projection_comm(Tij->device_ptr(), Tij->lattice().size(0), Tij->lattice().halo(),
                                       LF2MI_PROJ_SYMTENSOR_NGP, lf2mi_real_code(), nullptr),
                 "symtensorProjectionCICNGP_comm");
}

}  // namespace LATfield2

#endif
