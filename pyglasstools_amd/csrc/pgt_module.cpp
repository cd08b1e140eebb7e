// pgt_module.cpp -- pybind11 mirror of the four pyglasstools extension
// modules on the IK stress-field path.  One extension (_pgt) holds four
// submodules named after the reference's modules:
//   _pyglasstools   pyglasstools/module-pyglasstools.cpp:50-72
//   _potential      pyglasstools/potential/module-potential.cpp:23-35
//   _cgfunc         pyglasstools/irvingkirkwood/cgfunc/module-cgfunc.cpp:19-31
//   _irvingkirkwood pyglasstools/irvingkirkwood/module-irvingkirkwood.cpp:21-32
// Class names, constructor signatures, method names and argument meanings
// follow the reference.  All arithmetic of the path runs on the GPU through
// the C ABI of libpgt_ik (include/pgt_ik.h); this file only moves data.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgt_ik.h"

namespace py = pybind11;
using Vec3 = std::array<double, 3>;
using darray = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace {

[[noreturn]] void raise_pgt(int rc, const std::string &where)
{
    std::string msg = where + ": " + pgt_last_error();
    if (rc == PGT_ERR_ARG || rc == PGT_ERR_RANGE) throw py::value_error(msg);
    if (rc == PGT_ERR_UNSUPPORTED) throw std::runtime_error(msg);
    throw std::runtime_error(msg);
}

inline void check(int rc, const std::string &where)
{
    if (rc != PGT_OK) raise_pgt(rc, where);
}

Vec3 to_vec3(const darray &a, const char *what)
{
    auto r = a.unchecked();
    if (a.size() != 3) throw py::value_error(std::string(what) + " must have 3 components");
    const double *p = a.data();
    return {p[0], p[1], p[2]};
}

// N x 3 (or N x 2, padded with z = 0) array to a flat row-major vector
std::vector<double> rows3(const darray &a, size_t n, const char *what)
{
    if (a.ndim() != 2 || (size_t)a.shape(0) != n || a.shape(1) < 2 || a.shape(1) > 3)
        throw py::value_error(std::string("[ERROR]: Size of ") + what + " array mismatch with # of particles!");
    std::vector<double> out(n * 3, 0.0);
    const double *p = a.data();
    const size_t m = (size_t)a.shape(1);
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < m; ++k) out[3 * i + k] = p[i * m + k];
    return out;
}

std::vector<double> vec1(const darray &a, size_t n, const char *what)
{
    if ((size_t)a.size() != n)
        throw py::value_error(std::string("[ERROR]: Size of ") + what + " array mismatch with # of particles!");
    return std::vector<double>(a.data(), a.data() + n);
}

py::array_t<double> view3(double *p, py::handle base) { return py::array_t<double>({3}, {sizeof(double)}, p, base); }

}  // namespace

// ---------------------------------------------------------------------------
// _pyglasstools: SimBox, ParticleSystem, Calculator, Observable, Manager,
// Communicator, MPILogFile
// ---------------------------------------------------------------------------

// SimBox.cc:3-52: box [-L/2 - origin, L/2 - origin); in 2D the z size is set
// to 1 and z is non-periodic.
struct SimBox {
    int dim = 3;
    double origin[3] = {0, 0, 0};
    double boxsize[3] = {1, 1, 1};
    double lowerbound[3] = {0, 0, 0};
    double upperbound[3] = {0, 0, 0};
    double periodic[3] = {1, 1, 1};
    double vol = 1;

    SimBox(const Vec3 &L, const Vec3 &o, int d) { init(L, o, d); }
    void init(const Vec3 &L, const Vec3 &o, int d)
    {
        dim = d;
        for (int a = 0; a < 3; ++a) {
            origin[a] = o[a];
            boxsize[a] = L[a];
        }
        setBounds();
        for (int a = 0; a < 3; ++a) periodic[a] = 1;
        if (dim == 2) {
            boxsize[2] = 1.0;
            periodic[2] = 0;
            origin[2] = 0.0;
            vol = boxsize[0] * boxsize[1];
        } else {
            vol = boxsize[0] * boxsize[1] * boxsize[2];
        }
    }
    void setBounds()
    {
        for (int a = 0; a < 3; ++a) {
            upperbound[a] = 0.5 * boxsize[a] - origin[a];
            lowerbound[a] = -upperbound[a] - 2.0 * origin[a];
        }
    }
    // SimBox.h:50-69, including its single-precision rounding of the image index
    Vec3 minImage(const Vec3 &v) const
    {
        Vec3 w = v;
        for (int a = 2; a >= 0; --a)
            if (periodic[a]) {
                double img = rintf((float)(w[a] / boxsize[a]));
                w[a] -= boxsize[a] * img;
            }
        return w;
    }
};

// ParticleSystem.h:14-64: host copies of the per-particle arrays; the
// neighbour index is built on the GPU by the engine at each compute.
struct ParticleSystem {
    std::shared_ptr<SimBox> simbox;
    size_t n = 0;
    std::vector<double> diameter, mass, position, velocity, displacement;  // N, N, Nx3, Nx3, Nx3
    uint64_t version = 0;  // bumped on every change, so calculators re-upload

    ParticleSystem(std::shared_ptr<SimBox> box, unsigned int numparticles, const darray &diam, const darray &m,
                   const darray &pos, const darray &vel)
        : simbox(std::move(box))
    {
        setData(numparticles, diam, m, pos, vel);
    }
    void setData(unsigned int numparticles, const darray &diam, const darray &m, const darray &pos, const darray &vel)
    {
        n = numparticles;
        diameter = vec1(diam, n, "diameter");
        mass = vec1(m, n, "mass");
        position = rows3(pos, n, "position");
        velocity = rows3(vel, n, "velocity");
        displacement.assign(n * 3, 0.0);
        ++version;
    }
    py::array_t<double> getN3(const std::vector<double> &v) const
    {
        py::array_t<double> a({(py::ssize_t)n, (py::ssize_t)3});
        std::copy(v.begin(), v.end(), a.mutable_data());
        return a;
    }
    // ParticleSystem.cc:75-188 (setters check sizes, as the reference does)
    void setMass(const darray &m) { mass = vec1(m, n, "mass"); ++version; }
    void setDiameter(const darray &d) { diameter = vec1(d, n, "diameter"); ++version; }
    void setAtomPosition(const darray &p) { position = rows3(p, n, "position"); ++version; }
    void setAtomVelocity(const darray &v) { velocity = rows3(v, n, "velocity"); ++version; }
    void setDisplacement(const darray &u) { displacement = rows3(u, n, "displacement"); ++version; }
    void moveParticles(const darray &d)
    {
        auto v = rows3(d, n, "position");
        for (size_t k = 0; k < 3 * n; ++k) position[k] += v[k];
        ++version;
    }
    void displaceParticles()
    {
        for (size_t k = 0; k < 3 * n; ++k) position[k] += displacement[k];
        ++version;
    }
    double dmin(double d, int a) const
    {
        if (simbox->periodic[a] != 0.0) d -= simbox->boxsize[a] * std::nearbyint(d / simbox->boxsize[a]);
        return d;
    }
    // ParticleSystem.cc:171-193: particles within `radius` of a point, by
    // minimum image; returned in index order.
    std::vector<unsigned int> getNeighbors(const Vec3 &pt, double radius) const
    {
        std::vector<unsigned int> out;
        const double r2 = radius * radius;
        for (size_t i = 0; i < n; ++i) {
            double s = 0;
            for (int a = 0; a < 3; ++a) {
                double d = dmin(position[3 * i + a] - pt[a], a);
                s += d * d;
            }
            if (s < r2) out.push_back((unsigned)i);
        }
        return out;
    }
    std::vector<Vec3> getNeighborsDistance(int tag, double radius) const
    {
        if (tag < 0 || (size_t)tag >= n) throw py::value_error("particle tag out of range");
        std::vector<Vec3> out;
        const double r2 = radius * radius;
        for (size_t j = 0; j < n; ++j) {
            Vec3 d;
            double s = 0;
            for (int a = 0; a < 3; ++a) {
                d[a] = dmin(position[3 * j + a] - position[3 * (size_t)tag + a], a);
                s += d[a] * d[a];
            }
            if (s < r2) out.push_back(d);
        }
        return out;
    }
};

struct PairPotential;

// Calculator.h:28-56: max_rcut from the diameters of the system it is built
// with; setSystemData swaps the system and keeps max_rcut.
struct Calculator {
    std::shared_ptr<ParticleSystem> m_sysdata;
    std::shared_ptr<PairPotential> m_potential;
    double max_rcut = 0;
    Calculator(std::shared_ptr<ParticleSystem> sys, std::shared_ptr<PairPotential> pot);
    virtual ~Calculator() = default;
    void setSystemData(std::shared_ptr<ParticleSystem> sys) { m_sysdata = std::move(sys); }
};

// Observable.h:12-25
struct Observable {
    std::string name = "NONE", type = "SCALAR";
    bool islocal = true;
    int dim = 3;
    Observable() = default;
    Observable(std::string n, std::string t, bool l, int d) : name(std::move(n)), type(std::move(t)), islocal(l), dim(d) {}
    virtual ~Observable() = default;
};

// Manager.h:136-187: notice level from "-notice_level N" on sys.argv
struct Manager {
    unsigned notice_level = 1;
    Manager()
    {
        try {
            py::list argv = py::module_::import("sys").attr("argv");
            for (size_t k = 0; k + 1 < argv.size(); ++k) {
                std::string a = py::str(argv[k]);
                if (a == "-notice_level") {
                    std::string v = py::str(argv[k + 1]);
                    try {
                        notice_level = (unsigned)std::stoul(v);
                    } catch (...) {
                        throw std::domain_error("[ERROR] Invalid value for notice_level");
                    }
                }
            }
        } catch (py::error_already_set &) {
        }
    }
    void notice(unsigned level, const std::string &message)
    {
        if (level <= notice_level && level > 1) py::print("notice(" + std::to_string(level) + "): " + message);
    }
};

// A single-rank communicator with the surface of MPI::Communicator
// (MPIInterface.cpp:142-163).  The engine needs no MPI: the whole grid is one
// GPU's work.
struct Communicator {
    int getRank() const { return 0; }
    int getRankGlobal() const { return 0; }
    int getSizeGlobal() const { return 1; }
    int getNRanks() const { return 1; }
    int getPartition() const { return 0; }
    void splitPartitions(unsigned nrank)
    {
        if (nrank == 0) throw std::runtime_error("--nrank setting has to be > 0");
        if (nrank != 1) throw std::runtime_error("Invalid setting --nrank.");
    }
    void barrier() {}
    void abort(int) {}  // MPI::Communicator::mpiabort is a no-op on one rank (MPIInterface.cpp:131-140)
};

// MPIFile.h:23-56 on one rank: an append-mode file whose writes reach the
// file when write_shared returns.
struct MPILogFile {
    FILE *fp = nullptr;
    MPILogFile(std::shared_ptr<Communicator>, const std::string &filename)
    {
        fp = std::fopen(filename.c_str(), "ab");
        if (!fp) throw std::runtime_error("cannot open " + filename);
    }
    ~MPILogFile()
    {
        if (fp) std::fclose(fp);
    }
    void write_shared(const std::string &s)
    {
        if (!s.empty() && std::fwrite(s.data(), 1, s.size(), fp) != s.size())
            throw std::runtime_error("write failed");
        std::fflush(fp);
    }
    void write_ordered(const std::string &s) { write_shared(s); }
};

// ---------------------------------------------------------------------------
// _potential
// ---------------------------------------------------------------------------
// PairPotential.h:14-67 / :69-123.  `kind` < 0 is the parameter-only base
// class, which evaluates to zero as the reference's base class does.
struct PairPotential {
    int kind = -1;
    double scaled_rcut = 0;
    std::vector<double> params;
    PairPotential(const std::vector<double> &p) : params(p) {}
    PairPotential(double rc, const std::vector<double> &p) : scaled_rcut(rc), params(p) {}
    virtual ~PairPotential() = default;
    void setParams(double val, int num)
    {
        if (num < 0 || (size_t)num >= params.size()) throw py::index_error("parameter index out of range");
        params[num] = val;
    }
    std::vector<double> getParams() const { return params; }
    std::array<double, 3> eval(const Vec3 &rij, double di, double dj) const
    {
        if (kind < 0) return {0, 0, 0};
        double f = 0, en = 0, rc = 0;
        check(pgt_eval_pair_force(kind, params.data(), (int)params.size(), scaled_rcut, 1, rij.data(), &di, &dj, &f,
                                  &en, &rc),
              "pair potential");
        return {f, en, rc};
    }
    double getRcut(const Vec3 &rij, double di, double dj) const { return eval(rij, di, dj)[2]; }
    double getPairForceDivR(const Vec3 &rij, double di, double dj) const { return eval(rij, di, dj)[0]; }
    double getPairEnergy(const Vec3 &rij, double di, double dj) const { return eval(rij, di, dj)[1]; }
};

template <int K>
struct ShortRangePairPotential : PairPotential {
    ShortRangePairPotential(const std::vector<double> &p) : PairPotential(p) { kind = K; }
    ShortRangePairPotential(double rc, const std::vector<double> &p) : PairPotential(rc, p) { kind = K; }
};

Calculator::Calculator(std::shared_ptr<ParticleSystem> sys, std::shared_ptr<PairPotential> pot)
    : m_sysdata(std::move(sys)), m_potential(std::move(pot))
{
    double maxforce_rcut = m_potential->scaled_rcut;
    if (m_sysdata->diameter.empty()) throw py::value_error("the particle system has no particles");
    double maxdiameter = *std::max_element(m_sysdata->diameter.begin(), m_sysdata->diameter.end());
    maxforce_rcut *= maxdiameter;
    max_rcut = std::max(maxforce_rcut, maxdiameter);
}

// ---------------------------------------------------------------------------
// _cgfunc
// ---------------------------------------------------------------------------
// CoarseGrainFunction.h:14-134.  kind < 0: the base class (returns zeros).
struct CoarseGrainFunction {
    int kind = -1;
    int quad_mode = PGT_QUAD_FIXED;
    int order = 1;
    int maxiter = 1;
    double relerr = 1e-9, abserr = 1e-9;
    double cg_rcut = 0;
    explicit CoarseGrainFunction(double rc) : cg_rcut(rc) {}
    virtual ~CoarseGrainFunction() = default;
    void setRcut(double rc) { cg_rcut = rc; }
    double getRcut() const { return cg_rcut; }
    double eval(int mode, double s, const Vec3 &x, const Vec3 &ri, const Vec3 *rij) const
    {
        if (kind < 0) return 0.0;
        double out = 0;
        check(pgt_eval_cgfunc(kind, cg_rcut, quad_mode, order, maxiter, relerr, abserr, mode, 1, &s, x.data(), ri.data(),
                              rij ? rij->data() : nullptr, &out),
              "coarse-graining function");
        return out;
    }
    double getDeltaFunc(const Vec3 &x, const Vec3 &ri) const { return eval(0, 0.0, x, ri, nullptr); }
    double getObjFunc(double s, const Vec3 &x, const Vec3 &ri, const Vec3 &rij) const { return eval(1, s, x, ri, &rij); }
    double getBondFunc(const Vec3 &x, const Vec3 &ri, const Vec3 &rij) const { return eval(2, 0.0, x, ri, &rij); }
};

template <int K>
struct FixedPointCGFunc : CoarseGrainFunction {
    FixedPointCGFunc(int ord, double rc) : CoarseGrainFunction(rc)
    {
        kind = K;
        quad_mode = PGT_QUAD_FIXED;
        order = ord;
    }
};

template <int K>
struct AdaptiveCGFunc : CoarseGrainFunction {
    AdaptiveCGFunc(int neval, double rel, double abs_, double rc) : CoarseGrainFunction(rc)
    {
        kind = K;
        quad_mode = PGT_QUAD_ADAPTIVE;
        maxiter = neval;
        relerr = rel;
        abserr = abs_;
    }
};

// ---------------------------------------------------------------------------
// _irvingkirkwood: fields and the calculator
// ---------------------------------------------------------------------------
// CoarseGrainedField.h:8-46
struct CoarseGrainedField : Observable {
    std::vector<Vec3> gridpoints;
    CoarseGrainedField() = default;
    CoarseGrainedField(std::string n, std::string t, bool l, int d) : Observable(std::move(n), std::move(t), l, d) {}
    void addGridpoints(const darray &g)
    {
        if (g.ndim() != 2 || g.shape(1) != 3) throw py::value_error("grid points must be an (M, 3) array");
        gridpoints.resize((size_t)g.shape(0));
        const double *p = g.data();
        for (size_t i = 0; i < gridpoints.size(); ++i) gridpoints[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    }
};

enum class FieldKind { Virial, Elastic, Kinetic };

// ForceField / LocalField (CoarseGrainedField.h:50-216): one D x D tensor per
// grid point, stored flat with index a + D*b.
struct TensorField : CoarseGrainedField {
    FieldKind fkind;
    int D;
    size_t gridsize;
    std::vector<double> val;
    TensorField(FieldKind k, int d, std::string n, std::string t, bool l, int gs)
        : CoarseGrainedField(std::move(n), std::move(t), l, d), fkind(k), D(d)
    {
        if (gs < 0) throw py::value_error("gridsize must be non-negative");
        gridsize = (size_t)gs;
        val.assign(gridsize * D * D, 0.0);
    }
    void check_index(int index, int grid_id) const
    {
        if (grid_id < 0 || (size_t)grid_id >= gridsize) throw py::index_error("grid index out of range");
        if (index < 0 || index >= D * D) throw py::index_error("tensor index out of range");
    }
    // gettostring (CoarseGrainedField.h:160-183): default stream formatting
    // (6 significant digits) followed by a space
    std::string gettostring(int index, int grid_id) const
    {
        check_index(index, grid_id);
        std::ostringstream os;
        os << val[(size_t)grid_id * D * D + index] << " ";
        return os.str();
    }
    // gettostring for grid points 0..count-1 in one call (same formatting)
    std::vector<std::string> formatColumn(int index, int count) const
    {
        if (count < 0 || (size_t)count > gridsize) throw py::index_error("grid count out of range");
        if (index < 0 || index >= D * D) throw py::index_error("tensor index out of range");
        std::vector<std::string> out((size_t)count);
        std::ostringstream os;
        for (size_t g = 0; g < (size_t)count; ++g) {
            os.str(std::string());
            os.clear();
            os << val[g * D * D + index] << " ";
            out[g] = os.str();
        }
        return out;
    }
    void save(std::shared_ptr<MPILogFile> f, int index, int grid_id) const { f->write_shared(gettostring(index, grid_id)); }
    void clear(int grid_id)
    {
        if (grid_id < 0 || (size_t)grid_id >= gridsize) throw py::index_error("grid index out of range");
        std::fill(val.begin() + (size_t)grid_id * D * D, val.begin() + (size_t)(grid_id + 1) * D * D, 0.0);
    }
};

template <FieldKind K, int Dim>
struct FieldT : TensorField {
    FieldT(std::string n, std::string t, bool l, int gs) : TensorField(K, Dim, std::move(n), std::move(t), l, gs) {}
};

// IrvingKirkwood.h:8-71, with compute running on the GPU engine.
struct IrvingKirkwood : Calculator {
    std::shared_ptr<CoarseGrainFunction> m_cgfunc;
    std::shared_ptr<Communicator> m_comm;
    std::map<std::string, std::shared_ptr<TensorField>> m_observables;
    pgt_ik_engine *engine = nullptr;
    int engine_dim = 0;
    int precision = PGT_PREC_FP64;
    int threads_per_block = 0;
    pgt_ik_stats last_stats{};

    IrvingKirkwood(std::shared_ptr<ParticleSystem> sys, std::shared_ptr<PairPotential> pot,
                   std::shared_ptr<CoarseGrainFunction> cg, std::shared_ptr<Communicator> comm)
        : Calculator(std::move(sys), std::move(pot)), m_cgfunc(std::move(cg)), m_comm(std::move(comm))
    {
    }
    ~IrvingKirkwood() override
    {
        if (engine) pgt_ik_destroy(engine);
    }
    void addObservable(std::shared_ptr<TensorField> obs)
    {
        // std::map::insert keeps the first field registered under a name
        m_observables.insert({obs->name, obs});
    }
    void setPrecision(const std::string &p)
    {
        int v = p == "fp64" ? PGT_PREC_FP64 : p == "fp32" ? PGT_PREC_FP32 : -1;
        if (v < 0) throw py::value_error("precision must be 'fp64' or 'fp32'");
        if (v != precision && engine) {
            pgt_ik_destroy(engine);
            engine = nullptr;
        }
        precision = v;
    }
    void ensure_engine(int dim)
    {
        if (engine && engine_dim == dim) return;
        if (engine) {
            pgt_ik_destroy(engine);
            engine = nullptr;
        }
        if (m_potential->kind < 0) throw py::value_error("the calculator needs a concrete pair potential");
        if (m_cgfunc->kind < 0) throw py::value_error("the calculator needs a concrete coarse-graining function");
        pgt_ik_params p{};
        p.dim = dim;
        p.potential_kind = m_potential->kind;
        p.scaled_rcut = m_potential->scaled_rcut;
        p.n_pot_params = (int)std::min<size_t>(6, m_potential->params.size());
        for (int k = 0; k < p.n_pot_params; ++k) p.pot_params[k] = m_potential->params[k];
        p.cg_kind = m_cgfunc->kind;
        p.cg_rcut = m_cgfunc->cg_rcut;
        p.quad_mode = m_cgfunc->quad_mode;
        p.quad_order = m_cgfunc->order;
        p.adapt_limit = m_cgfunc->maxiter;
        p.adapt_epsrel = m_cgfunc->relerr;
        p.adapt_epsabs = m_cgfunc->abserr;
        p.precision = precision;
        p.max_rcut = max_rcut;
        check(pgt_ik_create(&p, &engine), "IrvingKirkwood");
        engine_dim = dim;
        if (threads_per_block) check(pgt_ik_set_tuning(engine, threads_per_block, 0, 0, 0), "IrvingKirkwood");
    }
    void compute(const darray &grid)
    {
        if (grid.ndim() != 2 || grid.shape(1) != 3) throw py::value_error("grid points must be an (M, 3) array");
        const size_t G = (size_t)grid.shape(0);
        // fields by kind; their tensor size fixes the output dimension
        std::shared_ptr<TensorField> fv, fe, fk;
        int dim = 0;
        for (auto &kv : m_observables) {
            auto &f = kv.second;
            if (f->gridsize < G) throw py::value_error("field '" + f->name + "' holds fewer grid points than compute was given");
            if (dim && f->D != dim) throw py::value_error("all fields of one calculator must have the same dimension");
            dim = f->D;
            if (f->fkind == FieldKind::Virial && !fv) fv = f;
            else if (f->fkind == FieldKind::Elastic && !fe) fe = f;
            else if (f->fkind == FieldKind::Kinetic && !fk) fk = f;
        }
        if (!dim || G == 0) return;
        ensure_engine(dim);
        const ParticleSystem &s = *m_sysdata;
        int per[3] = {s.simbox->periodic[0] != 0, s.simbox->periodic[1] != 0, s.simbox->periodic[2] != 0};
        check(pgt_ik_set_box(engine, s.simbox->boxsize, s.simbox->lowerbound, per), "IrvingKirkwood.compute");
        check(pgt_ik_set_particles(engine, (int64_t)s.n, s.position.data(), s.diameter.data(), s.mass.data(),
                                   s.velocity.data(), fe ? s.displacement.data() : nullptr),
              "IrvingKirkwood.compute");
        check(pgt_ik_set_grid(engine, (int64_t)G, grid.data()), "IrvingKirkwood.compute");
        unsigned fields = (fv ? PGT_FIELD_VIRIAL : 0) | (fe ? PGT_FIELD_ELASTIC : 0) | (fk ? PGT_FIELD_KINETIC : 0);
        {
            py::gil_scoped_release nogil;
            check(pgt_ik_compute(engine, fields, fv ? fv->val.data() : nullptr, fk ? fk->val.data() : nullptr,
                                 fe ? fe->val.data() : nullptr),
                  "IrvingKirkwood.compute");
        }
        check(pgt_ik_get_stats(engine, &last_stats), "IrvingKirkwood.compute");
    }
    void printDisplacement() const
    {
        for (size_t i = 0; i < m_sysdata->n; ++i)
            py::print(m_sysdata->displacement[3 * i], m_sysdata->displacement[3 * i + 1], (int)i);
    }
    py::dict getStats() const
    {
        py::dict d;
        d["n_particles"] = last_stats.n_particles;
        d["n_grid"] = last_stats.n_grid;
        d["n_ball"] = last_stats.n_ball;
        d["n_contrib"] = last_stats.n_contrib;
        d["bond_capacity"] = last_stats.bond_capacity;
        d["gather_ms"] = last_stats.gather_ms;
        d["total_ms"] = last_stats.total_ms;
        return d;
    }
};

template <class T>
void export_pot(py::module_ &m, const char *name)
{
    py::class_<T, PairPotential, std::shared_ptr<T>>(m, name)
        .def(py::init<std::vector<double>>())
        .def(py::init<double, std::vector<double>>());
}

template <class T>
void export_field(py::module_ &m, const char *name)
{
    py::class_<T, TensorField, std::shared_ptr<T>>(m, name).def(py::init<std::string, std::string, bool, int>());
}

PYBIND11_MODULE(_pgt, root)
{
    root.doc() = "MI355X IK stress-field engine: pybind11 surface of the pyglasstools IK path";
    root.def("device_count", &pgt_device_count);

    // ---- _pyglasstools ----
    auto m = root.def_submodule("_pyglasstools");
    py::class_<SimBox, std::shared_ptr<SimBox>>(m, "SimBox")
        .def(py::init([](double L, const darray &o, int d) {
                 return std::make_shared<SimBox>(Vec3{L, L, L}, to_vec3(o, "origin"), d);
             }),
             py::arg("boxsize"), py::arg("origin"), py::arg("dim"))
        .def(py::init([](const darray &L, const darray &o, int d) {
                 return std::make_shared<SimBox>(to_vec3(L, "boxsize"), to_vec3(o, "origin"), d);
             }),
             py::arg("boxsize"), py::arg("origin"), py::arg("dim"))
        .def("minImage", [](const SimBox &b, const darray &v) { return b.minImage(to_vec3(v, "vector")); })
        .def_readwrite("dim", &SimBox::dim)
        .def_readwrite("vol", &SimBox::vol)
        .def_property("origin", [](py::object self) { return view3(self.cast<SimBox &>().origin, self); },
                      [](SimBox &b, const darray &v) { auto x = to_vec3(v, "origin"); std::copy(x.begin(), x.end(), b.origin); })
        .def_property("boxsize", [](py::object self) { return view3(self.cast<SimBox &>().boxsize, self); },
                      [](SimBox &b, const darray &v) { auto x = to_vec3(v, "boxsize"); std::copy(x.begin(), x.end(), b.boxsize); })
        .def_property("upperbound", [](py::object self) { return view3(self.cast<SimBox &>().upperbound, self); },
                      [](SimBox &b, const darray &v) { auto x = to_vec3(v, "upperbound"); std::copy(x.begin(), x.end(), b.upperbound); })
        .def_property("lowerbound", [](py::object self) { return view3(self.cast<SimBox &>().lowerbound, self); },
                      [](SimBox &b, const darray &v) { auto x = to_vec3(v, "lowerbound"); std::copy(x.begin(), x.end(), b.lowerbound); })
        .def_property("periodic", [](py::object self) { return view3(self.cast<SimBox &>().periodic, self); },
                      [](SimBox &b, const darray &v) { auto x = to_vec3(v, "periodic"); std::copy(x.begin(), x.end(), b.periodic); });

    py::class_<ParticleSystem, std::shared_ptr<ParticleSystem>>(m, "ParticleSystem")
        .def(py::init<std::shared_ptr<SimBox>, unsigned int, const darray &, const darray &, const darray &, const darray &>())
        .def("getMass", [](const ParticleSystem &s) { return s.mass; })
        .def("setMass", &ParticleSystem::setMass)
        .def("getDiameter", [](const ParticleSystem &s) { return s.diameter; })
        .def("setDiameter", &ParticleSystem::setDiameter)
        .def("getDisplacement", [](const ParticleSystem &s) { return s.getN3(s.displacement); })
        .def("setDisplacement", &ParticleSystem::setDisplacement)
        .def("setAtomPosition", &ParticleSystem::setAtomPosition)
        .def("setAtomVelocity", &ParticleSystem::setAtomVelocity)
        .def("getAtomPosition", [](const ParticleSystem &s) { return s.getN3(s.position); })
        .def("getAtomVelocity", [](const ParticleSystem &s) { return s.getN3(s.velocity); })
        .def("moveParticles", &ParticleSystem::moveParticles)
        .def("moveParticles", &ParticleSystem::displaceParticles)
        .def("getNeighbors", [](const ParticleSystem &s, const darray &p, double r) { return s.getNeighbors(to_vec3(p, "point"), r); })
        .def("getNeighborsDistance", &ParticleSystem::getNeighborsDistance)
        .def("update", &ParticleSystem::setData)
        .def_property_readonly("simbox", [](const ParticleSystem &s) { return s.simbox; })
        .def_property_readonly("N", [](const ParticleSystem &s) { return s.n; });

    py::class_<Calculator, std::shared_ptr<Calculator>>(m, "Calculator")
        .def(py::init<std::shared_ptr<ParticleSystem>, std::shared_ptr<PairPotential>>())
        .def_readonly("max_rcut", &Calculator::max_rcut);
    py::class_<Observable, std::shared_ptr<Observable>>(m, "Observable")
        .def(py::init<std::string, std::string, bool, int>())
        .def_readwrite("name", &Observable::name)
        .def_readwrite("type", &Observable::type)
        .def_readwrite("islocal", &Observable::islocal)
        .def_readwrite("dim", &Observable::dim);
    py::class_<Manager, std::shared_ptr<Manager>>(m, "Manager").def(py::init<>()).def("notice", &Manager::notice);
    py::class_<Communicator, std::shared_ptr<Communicator>>(m, "Communicator")
        .def(py::init<>())
        .def("splitPartitions", &Communicator::splitPartitions)
        .def("getPartition", &Communicator::getPartition)
        .def("getNRanks", &Communicator::getNRanks)
        .def("getRank", &Communicator::getRank)
        .def("getRankGlobal", &Communicator::getRankGlobal)
        .def("getSizeGlobal", &Communicator::getSizeGlobal)
        .def("barrier", &Communicator::barrier)
        .def("abort", &Communicator::abort)
        .def("scatter_v", [](const Communicator &, py::list in, unsigned root) -> py::object {
            if (root != 0) throw py::value_error("root out of range");
            if (in.size() < 1) throw py::value_error("scatter_v needs one entry per rank");
            return py::module_::import("numpy").attr("array")(in[0], "float64");
        }, py::arg("in_values"), py::arg("root") = 0)
        .def("all_gather_v", [](const Communicator &, py::object v) {
            py::list l;
            l.append(v);
            return l;
        })
        .def("bcast", [](const Communicator &, py::object v, unsigned root) {
            if (root != 0) throw py::value_error("root out of range");
            return v;
        }, py::arg("val"), py::arg("root") = 0)
        .def("send", [](const Communicator &, bool, unsigned dest, unsigned) {
            throw py::value_error("send: rank " + std::to_string(dest) + " does not exist in a one-rank run");
        })
        .def("recv", [](const Communicator &, unsigned src, unsigned) -> bool {
            throw py::value_error("recv: rank " + std::to_string(src) + " does not exist in a one-rank run");
        });
    py::class_<MPILogFile, std::shared_ptr<MPILogFile>>(m, "MPILogFile")
        .def(py::init<std::shared_ptr<Communicator>, std::string>())
        .def("write_shared", &MPILogFile::write_shared)
        .def("write_ordered", &MPILogFile::write_ordered);

    // ---- _potential ----
    auto mp = root.def_submodule("_potential");
    auto v3 = [](const darray &a) { return to_vec3(a, "rij"); };
    py::class_<PairPotential, std::shared_ptr<PairPotential>>(mp, "PairPotential")
        .def(py::init<std::vector<double>>())
        .def(py::init<double, std::vector<double>>())
        .def("getRcut", [v3](const PairPotential &p, const darray &r, double di, double dj) { return p.getRcut(v3(r), di, dj); })
        .def("getPairForceDivR", [v3](const PairPotential &p, const darray &r, double di, double dj) { return p.getPairForceDivR(v3(r), di, dj); })
        .def("getPairEnergy", [v3](const PairPotential &p, const darray &r, double di, double dj) { return p.getPairEnergy(v3(r), di, dj); })
        .def("getParams", &PairPotential::getParams)
        .def("setParams", &PairPotential::setParams)
        .def_readwrite("scaled_rcut", &PairPotential::scaled_rcut);
    export_pot<ShortRangePairPotential<PGT_POT_POLY12>>(mp, "PairPotentialPoly12");
    export_pot<ShortRangePairPotential<PGT_POT_POLY18>>(mp, "PairPotentialPoly18");
    export_pot<ShortRangePairPotential<PGT_POT_POLY10>>(mp, "PairPotentialPoly10");
    export_pot<ShortRangePairPotential<PGT_POT_POLY106>>(mp, "PairPotentialPoly106");
    export_pot<ShortRangePairPotential<PGT_POT_POLYLJ>>(mp, "PairPotentialPolyLJ");

    // ---- _cgfunc ----
    auto mc = root.def_submodule("_cgfunc");
    auto vx = [](const darray &a) { return to_vec3(a, "vector"); };
    py::class_<CoarseGrainFunction, std::shared_ptr<CoarseGrainFunction>>(mc, "CoarseGrainFunction")
        .def(py::init<double>())
        .def("getRcut", &CoarseGrainFunction::getRcut)
        .def("setRcut", &CoarseGrainFunction::setRcut)
        .def("getDeltaFunc", [vx](const CoarseGrainFunction &c, const darray &x, const darray &ri) { return c.getDeltaFunc(vx(x), vx(ri)); })
        .def("getObjFunc", [vx](const CoarseGrainFunction &c, double s, const darray &x, const darray &ri, const darray &rij) {
            return c.getObjFunc(s, vx(x), vx(ri), vx(rij));
        })
        .def("getBondFunc", [vx](const CoarseGrainFunction &c, const darray &x, const darray &ri, const darray &rij) {
            return c.getBondFunc(vx(x), vx(ri), vx(rij));
        });
    py::class_<FixedPointCGFunc<PGT_CG_OCTIC>, CoarseGrainFunction, std::shared_ptr<FixedPointCGFunc<PGT_CG_OCTIC>>>(mc, "CGFuncOcticFixedPoint")
        .def(py::init<int, double>());
    py::class_<AdaptiveCGFunc<PGT_CG_OCTIC>, CoarseGrainFunction, std::shared_ptr<AdaptiveCGFunc<PGT_CG_OCTIC>>>(mc, "CGFuncOcticAdaptive")
        .def(py::init<int, double, double, double>());
    py::class_<FixedPointCGFunc<PGT_CG_RECT>, CoarseGrainFunction, std::shared_ptr<FixedPointCGFunc<PGT_CG_RECT>>>(mc, "CGFuncRectFixedPoint")
        .def(py::init<int, double>());
    py::class_<AdaptiveCGFunc<PGT_CG_RECT>, CoarseGrainFunction, std::shared_ptr<AdaptiveCGFunc<PGT_CG_RECT>>>(mc, "CGFuncRectAdaptive")
        .def(py::init<int, double, double, double>());
    py::class_<FixedPointCGFunc<PGT_CG_MOLLIFIER>, CoarseGrainFunction, std::shared_ptr<FixedPointCGFunc<PGT_CG_MOLLIFIER>>>(mc, "CGFuncMollifierFixedPoint")
        .def(py::init<int, double>());
    py::class_<AdaptiveCGFunc<PGT_CG_MOLLIFIER>, CoarseGrainFunction, std::shared_ptr<AdaptiveCGFunc<PGT_CG_MOLLIFIER>>>(mc, "CGFuncMollifierAdaptive")
        .def(py::init<int, double, double, double>());

    // ---- _irvingkirkwood ----
    auto mi = root.def_submodule("_irvingkirkwood");
    py::class_<IrvingKirkwood, Calculator, std::shared_ptr<IrvingKirkwood>>(mi, "IrvingKirkwood")
        .def(py::init<std::shared_ptr<ParticleSystem>, std::shared_ptr<PairPotential>, std::shared_ptr<CoarseGrainFunction>,
                      std::shared_ptr<Communicator>>())
        .def("compute", &IrvingKirkwood::compute)
        .def("setSystemData", &IrvingKirkwood::setSystemData)
        .def("addObservable", &IrvingKirkwood::addObservable)
        .def("printDisplacement", &IrvingKirkwood::printDisplacement)
        // additions beyond the reference surface
        .def("setPrecision", &IrvingKirkwood::setPrecision)
        .def("setThreadsPerBlock", [](IrvingKirkwood &k, int b) { k.threads_per_block = b; if (k.engine) check(pgt_ik_set_tuning(k.engine, b, 0, 0, 0), "setThreadsPerBlock"); })
        .def("getStats", &IrvingKirkwood::getStats)
        .def_readonly("max_rcut", &IrvingKirkwood::max_rcut);
    py::class_<CoarseGrainedField, Observable, std::shared_ptr<CoarseGrainedField>>(mi, "CoarseGrainedField");
    py::class_<TensorField, CoarseGrainedField, std::shared_ptr<TensorField>>(mi, "TensorField")
        .def("addGridpoints", &TensorField::addGridpoints)
        .def("gettostring", &TensorField::gettostring)
        .def("formatColumn", &TensorField::formatColumn)
        .def("save", &TensorField::save)
        .def("clear", &TensorField::clear)
        // numpy view of the (gridsize, D, D) values; [g, b, a] holds flat index a + D*b
        .def_property_readonly("values", [](py::object self) {
            auto &f = self.cast<TensorField &>();
            return py::array_t<double>({(py::ssize_t)f.gridsize, (py::ssize_t)f.D, (py::ssize_t)f.D}, f.val.data(), self);
        });
    export_field<FieldT<FieldKind::Virial, 2>>(mi, "VirialStressField2D");
    export_field<FieldT<FieldKind::Virial, 3>>(mi, "VirialStressField3D");
    export_field<FieldT<FieldKind::Elastic, 2>>(mi, "ElasticVirialStressField2D");
    export_field<FieldT<FieldKind::Elastic, 3>>(mi, "ElasticVirialStressField3D");
    export_field<FieldT<FieldKind::Kinetic, 2>>(mi, "KineticStressField2D");
    export_field<FieldT<FieldKind::Kinetic, 3>>(mi, "KineticStressField3D");
}
