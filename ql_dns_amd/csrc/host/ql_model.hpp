// ql_model.hpp -- C++ host mirror of the reference's Model / Integrator /
// solution interface, implemented on the C ABI of include/ql_dns.h.
//
// The classes keep the reference's names, virtual signatures and argument
// meaning so that its driver loop (main.cpp:108-154) can be pointed at the GPU
// path:
//   Model        Models/Model.h:24-112      (rhs, linearOPs_Init, sizes, remap,
//                                            noise, energy, CFL)
//   Integrator   Integrators/Integrator.h:20-28 (Step, mean_time_step)
//   solution     solution.h:23-49           (pLin, pMF)
//   GpuModel     stands in for MHD_FullQlin / MHD_BQlin (main.cpp:58-74)
//   GpuIntegrator stands in for RK3CN / RK3CN_Elin (main.cpp:95)
// Eigen is not used; dcmplxVec / doubVec are small owning arrays with the
// element access the reference's call sites use (operator(), size(), data()).
//
// State ownership: the device copy (owned by the GpuModel's ql_ctx) is
// authoritative while stepping.  A solution bound to a GpuModel mirrors it on
// the host: pLin / pMF first download a stale mirror, then mark it modified,
// and the next device operation uploads it.  Errors: the reference aborts the
// MPI job (General_Definitions.h:68-72); here failures throw qlgpu::Error with
// the C ABI's message.
#pragma once

#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ql_dns.h"

namespace qlgpu {

typedef std::complex<double> dcmplx;

template <typename T>
class Vec {
public:
    Vec() = default;
    explicit Vec(int n) : v_(n) {}
    int size() const { return (int)v_.size(); }
    T *data() { return v_.data(); }
    const T *data() const { return v_.data(); }
    T &operator()(int i) { return v_[i]; }
    const T &operator()(int i) const { return v_[i]; }
    T &operator[](int i) { return v_[i]; }
    const T &operator[](int i) const { return v_[i]; }
    void setZero() { for (auto &x : v_) x = T(0); }

private:
    std::vector<T> v_;
};
typedef Vec<dcmplx> dcmplxVec;
typedef Vec<double> doubVec;

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
    int code;
};

// What the reference reads from its input file for this path
// (Auxiliary/Input_parameters.cpp:68-138) plus the model/integrator choice.
struct Inputs {
    std::string equations_to_use = "MHD_FullQlin";  // or "MHD_BQlin"
    std::string integrator = "RK3CN_Elin";          // main.cpp:95; or "RK3CN"
    int NXY[2] = {16, 16};
    int NZ = 16;
    double L[3] = {6.283185307179586, 6.283185307179586, 6.283185307179586};
    double dt = 0.33333333333333;  // < 0: variable steps of at most -dt
    double CFL = 1.0;              // Input_parameters.cpp:77
    double t_start = 0.0, t_final = 10.0;
    double nu = 0.01, eta = 0.01, q = 1.5, omega = 1.0, B0z = 0.0;
    int viscosity_order = 2;
    double f_noise = 0.0;
    double noise_range_low = 0.0, noise_range_high = 1e16;
    bool drive_only_velocityQ = false, drive_only_magneticQ = false;
    unsigned long long noise_seed = 0x5EED0004ull;  // the device generator's key
    bool remapQ = true, QuasiLinearQ = true;
    int device = 0;
};

class solution;
class TimeVariables;
class Model;

// Auxiliary/Kdata.h:17-34: the wavenumber tables, k stored as i*k as the
// reference stores them, for callers that read Model::K (FullSave_Load.cpp:110,
// 134, 196, 233, 291, 344).  One device holds every mode, so the local and
// full index tables are the same.  kx is the current (remapped) kx0 of each
// mode; the model refreshes it after each remap, and a value the caller
// writes (a restart, FullSave_Load.cpp:344) reaches the device at the next
// device operation, as solution::pLin does.
class Kdata {
public:
    explicit Kdata(const Model &mod);
    Kdata(const Kdata &) = delete;
    Kdata &operator=(const Kdata &) = delete;
    dcmplx *kx = nullptr, *ky = nullptr;  // [nxy]
    int *ky_index = nullptr, *kx_index = nullptr;
    int *ky_index_full = nullptr, *kx_index_full = nullptr;
    dcmplxVec kz;                            // [nz], slot order (0..nzl, -nzl..-1)
    doubVec kz2;                             // -kz^2
    doubVec *lap2 = nullptr, *ilap2 = nullptr;  // per ky index: -ky^2 + kz2 and 1/that (1 at ky = kz = 0)
    double kxLfac = 0, kyLfac = 0, kzLfac = 0;

private:
    std::vector<dcmplx> kx_, ky_;
    std::vector<int> kxi_, kyi_;
    std::vector<doubVec> lap2_, ilap2_;
};

// Models/Model.h:24-112
class Model {
public:
    Model(int NZ, const int NXY[2], const double *L) : L_{L[0], L[1], L[2]}
    {
        NZ_ = NZ;
        Nxy_[0] = NXY[0];
        Nxy_[1] = NXY[1] / 2;
        nz_dealiased_ = 1 + 2 * int(NZ_ / 3.0);
        nxy_full_ = (Nxy_[0] - 1) * Nxy_[1];
    }
    virtual ~Model() {}
    virtual void rhs(const double t, const double dt_lin, solution *SolIn, solution *SolOut,
                     doubVec **linOpFluct) = 0;
    virtual void linearOPs_Init(double t0, doubVec **linOpFluct, doubVec *linOpMF) = 0;
    int Dimxy_full() const { return nxy_full_; }
    int NZ() const { return nz_dealiased_; }
    int NZfull() const { return NZ_; }
    virtual int Dimxy() const = 0;
    virtual int num_MFs() const = 0;
    virtual int num_Lin() const = 0;
    double L(int ind) const { return L_[ind]; }
    int N(int ind) const { return ind == 0 ? Nxy_[0] : ind == 1 ? Nxy_[1] * 2 : ind == 2 ? NZ_ : 0; }
    virtual void ShearingBox_Remap(double qt, solution *sol) = 0;
    virtual void DrivingNoise(double t, double dt, solution *sol) = 0;
    virtual void ChangeNoiseRange(double fnoise, double kmin, double kmax) = 0;
    virtual int num_reynolds_saves() = 0;
    virtual void Calc_Energy_AM_Diss(TimeVariables *tv, double t, const solution *sol) = 0;
    virtual double Calculate_CFL(const solution *sol) = 0;
    Kdata *K = nullptr;  // Model.h:74

protected:
    int Nxy_[2];
    int NZ_;
    int nz_dealiased_;
    int nxy_full_;
    double L_[3];
};

// Holds the per-save quantities Calc_Energy_AM_Diss fills
// (Auxiliary/TimeVariables.h: current_energy/AM/diss/reynolds, etaK_times_kmax)
// and records each save in memory.
class TimeVariables {
public:
    double *current_energy() { return energy_; }
    double *current_AM() { return AM_; }
    double *current_diss() { return diss_; }
    // num_reynolds_saves values (TimeVariables.cpp:30-40)
    double *current_reynolds(int n = 4)
    {
        if ((int)reynolds_.size() < n) reynolds_.resize(n, 0.0);
        return reynolds_.data();
    }
    double etaK_times_kmax[2] = {0, 0};
    void Save_Data(double t);
    struct Record {
        double t, energy[4], AM[4], diss[4];
        std::vector<double> reynolds;
    };
    const std::vector<Record> &records() const { return records_; }

private:
    double energy_[4] = {0, 0, 0, 0}, AM_[4] = {0, 0, 0, 0}, diss_[4] = {0, 0, 0, 0};
    std::vector<double> reynolds_ = std::vector<double>(4, 0.0);
    std::vector<Record> records_;
};

// The GPU path behind the Model interface (MHD_FullQlin, MHD_BQlin or HD_fullU).
class GpuModel : public Model {
public:
    explicit GpuModel(const Inputs &sp);
    ~GpuModel() override;
    GpuModel(const GpuModel &) = delete;
    GpuModel &operator=(const GpuModel &) = delete;

    void rhs(const double t, const double dt_lin, solution *SolIn, solution *SolOut, doubVec **linOpFluct) override;
    void linearOPs_Init(double t0, doubVec **linOpFluct, doubVec *linOpMF) override;
    int Dimxy() const override { return dims_[0]; }
    int num_MFs() const override { return dims_[2]; }
    int num_Lin() const override { return dims_[3]; }
    void ShearingBox_Remap(double qt, solution *sol) override;
    void DrivingNoise(double t, double dt, solution *sol) override;
    void ChangeNoiseRange(double fnoise, double kmin, double kmax) override;
    // HD_fullU saves the full Ux, Uy stress profiles (HD_fullU.cpp:50-52)
    int num_reynolds_saves() override { return hd_ ? 2 * NZfull() : 4; }
    void Calc_Energy_AM_Diss(TimeVariables *tv, double t, const solution *sol) override;
    double Calculate_CFL(const solution *sol) override;

    // current (possibly remapped) Im(kx) per mode from the device
    std::vector<double> kx() const;
    // sets every mode's kx0 (K->kx and the device)
    void set_kx(const double *kx0);
    // K->kx written by the caller since the device last saw it: upload it
    void push_kx();
    ql_ctx *ctx() const { return ctx_; }
    const Inputs &inputs() const { return sp_; }
    void check(int rc) const;
    // the solution whose contents are on the device (nullptr: none yet)
    solution *resident = nullptr;

private:
    Inputs sp_;
    ql_ctx *ctx_ = nullptr;
    int dims_[6] = {0, 0, 0, 0, 0, 0};
    double f_noise_ = 0.0;
    ql_noise noise_{};
    bool hd_ = false;
    Kdata *k_ = nullptr;             // owned; K points at it
    std::vector<double> kx_device_;  // kx0 as the device holds it
    void pull_kx();
};

// solution.h:23-49
class solution {
public:
    explicit solution(Model *eqs);
    dcmplxVec *pLin(int i, int Vnum) const
    {
        pull();
        dirty_ = true;
        return &lin_[(size_t)i * nLin_ + Vnum];
    }
    dcmplxVec *pMF(int i) const
    {
        pull();
        dirty_ = true;
        return &mf_[i];
    }
    // solution.cpp:169-203: mean fields real, bounded and finite, as a device
    // reduction; throws Error (QL_ERR_STATE) where the reference aborts
    void Check_Solution();
    // device <-> host mirror management (used by GpuModel / GpuIntegrator)
    void push();              // host -> device if the host copy was modified
    void pull() const;        // device -> host if the host copy is stale
    void mark_device_newer() { stale_ = true; }
    void pack(std::vector<dcmplx> &lin, std::vector<dcmplx> &mf) const;
    void unpack(const std::vector<dcmplx> &lin, const std::vector<dcmplx> &mf);

private:
    GpuModel *gm_;
    int nxy_, nLin_, nMF_, nz_;
    mutable std::vector<dcmplxVec> lin_;
    mutable std::vector<dcmplxVec> mf_;
    mutable bool stale_ = false, dirty_ = true;
};

// Integrators/Integrator.h:20-28
class Integrator {
public:
    virtual ~Integrator() {}
    virtual double Step(double t, solution *sol) = 0;
    virtual double mean_time_step() const = 0;
};

// RK3CN / RK3CN_Elin on the device (chosen by Inputs::integrator)
class GpuIntegrator : public Integrator {
public:
    GpuIntegrator(double t0, const Inputs &SP, GpuModel *mod);
    double Step(double t, solution *sol) override;
    double mean_time_step() const override { return step_count_ ? dt_mean_ / step_count_ : 0.0; }

private:
    GpuModel *model_;
    double dt_, dt_mean_ = 0.0;
    int step_count_ = 0;
};

}  // namespace qlgpu
