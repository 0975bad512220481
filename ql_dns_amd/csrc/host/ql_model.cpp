// ql_model.cpp -- the C++ Model / Integrator / solution mirror on the C ABI.
#include "ql_model.hpp"

#include <algorithm>
#include <cmath>

namespace qlgpu {

namespace {

int model_id(const std::string &s)
{
    if (s == "MHD_FullQlin") return QL_MODEL_MHD_FULLQLIN;
    if (s == "MHD_BQlin") return QL_MODEL_MHD_BQLIN;
    if (s == "HD_fullU") return QL_MODEL_HD_FULLU;
    throw Error(QL_ERR_ARG, "model '" + s + "' is not available on the GPU path (MHD_FullQlin, MHD_BQlin, HD_fullU)");
}

int integrator_id(const std::string &s)
{
    if (s == "RK3CN_Elin") return QL_INT_RK3CN_ELIN;
    if (s == "RK3CN") return QL_INT_RK3CN;
    throw Error(QL_ERR_ARG, "integrator '" + s + "' is not available on the GPU path (RK3CN, RK3CN_Elin)");
}


}  // namespace

void TimeVariables::Save_Data(double t)
{
    Record r;
    r.t = t;
    for (int k = 0; k < 4; ++k) {
        r.energy[k] = energy_[k];
        r.AM[k] = AM_[k];
        r.diss[k] = diss_[k];
    }
    r.reynolds = reynolds_;
    records_.push_back(r);
}

// the tables of Kdata.cpp:28-125 for one device holding every mode
Kdata::Kdata(const Model &mod)
{
    kxLfac = 2 * M_PI / mod.L(0);
    kyLfac = 2 * M_PI / mod.L(1);
    kzLfac = 2 * M_PI / mod.L(2);
    const int ny = mod.N(1) / 2, nx = mod.N(0) / 2, nxy = mod.Dimxy_full();
    kx_.resize(nxy);
    ky_.resize(nxy);
    kxi_.resize(nxy);
    kyi_.resize(nxy);
    for (int i = 0; i < 2 * nx; ++i) {
        if (i == nx) continue;  // the Nyquist kx is left out
        const int ix = i < nx ? i : -2 * nx + i, row = i < nx ? i : i - 1;
        for (int j = 0; j < ny; ++j) {
            kx_[j + ny * row] = dcmplx(0, ix * kxLfac);
            ky_[j + ny * row] = dcmplx(0, j * kyLfac);
            kxi_[j + ny * row] = ix;
            kyi_[j + ny * row] = j;
        }
    }
    kx = kx_.data();
    ky = ky_.data();
    kx_index = kx_index_full = kxi_.data();
    ky_index = ky_index_full = kyi_.data();
    const int nz = mod.NZ(), nzl = (nz - 1) / 2;
    kz = dcmplxVec(nz);
    kz2 = doubVec(nz);
    for (int i = 0; i < nz; ++i) {
        kz(i) = dcmplx(0, (i <= nzl ? i : -2 * nzl + i - 1) * kzLfac);
        kz2(i) = (kz(i) * kz(i)).real();
    }
    lap2_.assign(ny, doubVec(nz));
    ilap2_.assign(ny, doubVec(nz));
    for (int j = 0; j < ny; ++j) {
        const double kyr = ky_[j].imag();
        for (int m = 0; m < nz; ++m) {
            lap2_[j](m) = -kyr * kyr + kz2(m);
            ilap2_[j](m) = 1 / lap2_[j](m);
        }
        if (kyr == 0) {
            lap2_[j](0) = 0;
            ilap2_[j](0) = 1;
        }
    }
    lap2 = lap2_.data();
    ilap2 = ilap2_.data();
}

GpuModel::GpuModel(const Inputs &sp) : Model(sp.NZ, sp.NXY, sp.L), sp_(sp), f_noise_(sp.f_noise)
{
    ql_params p{};
    p.model = model_id(sp.equations_to_use);
    hd_ = p.model == QL_MODEL_HD_FULLU;
    p.integrator = integrator_id(sp.integrator);
    p.precision = QL_PREC_FP64;
    p.NX = sp.NXY[0];
    p.NY = sp.NXY[1];
    p.NZ = sp.NZ;
    for (int i = 0; i < 3; ++i) p.L[i] = sp.L[i];
    p.q = sp.q;
    p.omega = sp.omega;
    p.nu = sp.nu;
    p.eta = sp.eta;
    p.B0z = sp.B0z;
    p.viscosity_order = sp.viscosity_order;
    p.quasilinear = sp.QuasiLinearQ ? 1 : 0;
    p.remap = sp.remapQ ? 1 : 0;
    p.dt = sp.dt;
    p.CFL = sp.CFL;
    p.device = sp.device;
    check(ql_create(&p, &ctx_));
    check(ql_dims(ctx_, dims_));
    // Input_parameters.cpp:107-111 (MHD_FullQlin.cpp:18-22, 47-52)
    noise_ = ql_noise{sp.f_noise, sp.noise_range_low, sp.noise_range_high, sp.drive_only_velocityQ ? 1 : 0,
                      sp.drive_only_magneticQ ? 1 : 0, sp.noise_seed};
    if (sp.f_noise != 0.0) check(ql_set_noise(ctx_, &noise_));
    k_ = new Kdata(*this);
    K = k_;
    kx_device_.resize(Dimxy());
    pull_kx();
}

GpuModel::~GpuModel()
{
    delete k_;
    if (ctx_) ql_destroy(ctx_);
}

void GpuModel::pull_kx()
{
    check(ql_get_state(ctx_, nullptr, nullptr, kx_device_.data()));
    for (int i = 0; i < Dimxy(); ++i) K->kx[i] = dcmplx(0, kx_device_[i]);
}

void GpuModel::push_kx()
{
    bool changed = false;
    for (int i = 0; i < Dimxy() && !changed; ++i) changed = K->kx[i].imag() != kx_device_[i];
    if (!changed) return;
    for (int i = 0; i < Dimxy(); ++i) kx_device_[i] = K->kx[i].imag();
    check(ql_set_state(ctx_, nullptr, nullptr, kx_device_.data()));
}

void GpuModel::set_kx(const double *kx0)
{
    for (int i = 0; i < Dimxy(); ++i) K->kx[i] = dcmplx(0, kx0[i]);
    push_kx();
}

void GpuModel::check(int rc) const
{
    if (rc != QL_OK) throw Error(rc, ql_last_error());
}

void GpuModel::rhs(const double t, const double dt_lin, solution *SolIn, solution *SolOut, doubVec **linOpFluct)
{
    push_kx();
    SolIn->push();
    const int nxy = Dimxy(), nz = NZ(), nMF = num_MFs(), nL = num_Lin();
    std::vector<dcmplx> out((size_t)nxy * nL * nz), omf((size_t)nMF * nz);
    std::vector<double> L((size_t)nxy * nL * nz);
    check(ql_rhs(ctx_, t, dt_lin, reinterpret_cast<double *>(out.data()), reinterpret_cast<double *>(omf.data()),
                 L.data()));
    if (SolOut) SolOut->unpack(out, omf);
    if (linOpFluct)
        for (int i = 0; i < nxy; ++i)
            for (int j = 0; j < nL; ++j)
                for (int m = 0; m < nz; ++m) linOpFluct[i][j](m) = L[((size_t)i * nL + j) * nz + m];
}

void GpuModel::linearOPs_Init(double t0, doubVec **linOpFluct, doubVec *linOpMF)
{
    push_kx();
    const int nxy = Dimxy(), nz = NZ(), nMF = num_MFs(), nL = num_Lin();
    std::vector<double> L((size_t)nxy * nL * nz), Lmf((size_t)nMF * nz);
    check(ql_linear_ops_init(ctx_, t0, L.data(), Lmf.data()));
    if (linOpFluct)
        for (int i = 0; i < nxy; ++i)
            for (int j = 0; j < nL; ++j)
                for (int m = 0; m < nz; ++m) linOpFluct[i][j](m) = L[((size_t)i * nL + j) * nz + m];
    if (linOpMF)
        for (int f = 0; f < nMF; ++f)
            for (int m = 0; m < nz; ++m) linOpMF[f](m) = Lmf[(size_t)f * nz + m];
}

void GpuModel::ShearingBox_Remap(double qt, solution *sol)
{
    push_kx();
    sol->push();
    check(ql_remap(ctx_, qt));
    sol->mark_device_newer();
    pull_kx();  // K->kx of the remapped modes (MHD_FullQlin.cpp:474)
}

void GpuModel::DrivingNoise(double t, double dt, solution *sol)
{
    // MHD_FullQlin.cpp:786-1053 / MHD_BQlin.cpp:558-740 on the device
    push_kx();
    sol->push();
    check(ql_driving_noise(ctx_, t, dt));
    sol->mark_device_newer();
}

void GpuModel::ChangeNoiseRange(double fnoise, double kmin, double kmax)
{
    // MHD_FullQlin.h:46-50: amplitude and |k| range; the flags and seed stay
    f_noise_ = fnoise;
    noise_.f_noise = fnoise;
    noise_.range_low = kmin;
    noise_.range_high = kmax;
    check(ql_set_noise(ctx_, &noise_));
}

void GpuModel::Calc_Energy_AM_Diss(TimeVariables *tv, double t, const solution *sol)
{
    push_kx();
    const_cast<solution *>(sol)->push();
    double out[18];
    check(ql_diagnostics(ctx_, t, out));
    if (hd_) {
        // HD_fullU.cpp:466-535: (MU, u) pairs and the full stress profiles
        const int ix[2] = {0, 2};
        for (int k = 0; k < 4; ++k) {
            tv->current_energy()[k] = k < 2 ? out[ix[k]] : 0.0;
            tv->current_AM()[k] = k < 2 ? out[4 + ix[k]] : 0.0;
            tv->current_diss()[k] = k < 2 ? out[8 + ix[k]] : 0.0;
        }
        check(ql_drive_real(ctx_, tv->current_reynolds(num_reynolds_saves())));
        tv->etaK_times_kmax[0] = out[16];
        tv->etaK_times_kmax[1] = 0.0;
    } else {
        for (int k = 0; k < 4; ++k) {
            tv->current_energy()[k] = out[k];
            tv->current_AM()[k] = out[4 + k];
            tv->current_diss()[k] = out[8 + k];
            tv->current_reynolds()[k] = out[12 + k];
        }
        tv->etaK_times_kmax[0] = out[16];
        tv->etaK_times_kmax[1] = out[17];
    }
    tv->Save_Data(t);
}

double GpuModel::Calculate_CFL(const solution *sol)
{
    // MHD_FullQlin.cpp:769-780 / MHD_BQlin.cpp:540-552 on the maxima the
    // device records: profiles of the last rhs (and MHD_BQlin's current Bx)
    push_kx();
    const_cast<solution *>(sol)->push();
    double v = 0.0;
    check(ql_calculate_cfl(ctx_, &v));
    return v;
}

std::vector<double> GpuModel::kx() const
{
    std::vector<double> k(Dimxy());
    check(ql_get_state(ctx_, nullptr, nullptr, k.data()));
    return k;
}

solution::solution(Model *eqs)
{
    gm_ = dynamic_cast<GpuModel *>(eqs);
    if (!gm_) throw Error(QL_ERR_ARG, "solution: the model is not a GpuModel");
    nxy_ = gm_->Dimxy();
    nLin_ = gm_->num_Lin();
    nMF_ = gm_->num_MFs();
    nz_ = gm_->NZ();
    lin_.assign((size_t)nxy_ * nLin_, dcmplxVec(nz_));
    mf_.assign(nMF_, dcmplxVec(nz_));
}

void solution::pack(std::vector<dcmplx> &lin, std::vector<dcmplx> &mf) const
{
    pull();  // the device copy may be the newer one
    lin.resize((size_t)nxy_ * nLin_ * nz_);
    mf.resize((size_t)nMF_ * nz_);
    for (size_t k = 0; k < lin_.size(); ++k) std::copy(lin_[k].data(), lin_[k].data() + nz_, lin.data() + k * nz_);
    for (int f = 0; f < nMF_; ++f) std::copy(mf_[f].data(), mf_[f].data() + nz_, mf.data() + (size_t)f * nz_);
}

void solution::unpack(const std::vector<dcmplx> &lin, const std::vector<dcmplx> &mf)
{
    for (size_t k = 0; k < lin_.size(); ++k) std::copy(lin.data() + k * nz_, lin.data() + (k + 1) * nz_, lin_[k].data());
    for (int f = 0; f < nMF_; ++f) std::copy(mf.data() + (size_t)f * nz_, mf.data() + (size_t)(f + 1) * nz_, mf_[f].data());
    if (gm_->resident == this) gm_->resident = nullptr;  // host copy is now the newer one
    stale_ = false;
    dirty_ = true;
}

void solution::push()
{
    if (gm_->resident != this) {
        if (gm_->resident) gm_->resident->pull();  // keep the previous resident's data
        gm_->resident = this;
        dirty_ = true;
    }
    if (!dirty_) return;
    std::vector<dcmplx> lin, mf;
    pack(lin, mf);
    gm_->check(ql_set_state(gm_->ctx(), reinterpret_cast<const double *>(lin.data()),
                            reinterpret_cast<const double *>(mf.data()), nullptr));
    dirty_ = false;
    stale_ = false;
}

void solution::Check_Solution()
{
    // solution.cpp:169-203; the reference aborts where this throws
    push();
    gm_->check(ql_check_solution(gm_->ctx(), nullptr, nullptr));
}

void solution::pull() const
{
    if (gm_->resident != this || !stale_) return;
    std::vector<dcmplx> lin((size_t)nxy_ * nLin_ * nz_), mf((size_t)nMF_ * nz_);
    gm_->check(ql_get_state(gm_->ctx(), reinterpret_cast<double *>(lin.data()), reinterpret_cast<double *>(mf.data()),
                            nullptr));
    for (size_t k = 0; k < lin_.size(); ++k) std::copy(lin.data() + k * nz_, lin.data() + (k + 1) * nz_, lin_[k].data());
    for (int f = 0; f < nMF_; ++f) std::copy(mf.data() + (size_t)f * nz_, mf.data() + (size_t)(f + 1) * nz_, mf_[f].data());
    stale_ = false;
}

GpuIntegrator::GpuIntegrator(double t0, const Inputs &SP, GpuModel *mod) : model_(mod), dt_(SP.dt)
{
    // SP.dt < 0: variable steps (RK3CN.cpp:20-30), chosen by the device in ql_step
    // RK3CN.cpp:57 / RK3CN_Elin.cpp:257
    model_->linearOPs_Init(t0, nullptr, nullptr);
}

double GpuIntegrator::Step(double t, solution *sol)
{
    model_->push_kx();
    sol->push();
    double dt = 0.0;
    model_->check(ql_step(model_->ctx(), t, &dt));
    sol->mark_device_newer();
    dt_mean_ += dt;
    ++step_count_;
    return dt;
}

}  // namespace qlgpu
