// ql_io.cpp -- the reference's on-disk formats (include/ql_io.h): the HDF5
// full-solution dump / restart of Auxiliary/FullSave_Load.cpp and the binary
// time series of Auxiliary/TimeVariables.cpp, on host buffers; plus the
// reference-shaped C++ classes over the GPU model (ql_io.hpp).
#include "ql_io.hpp"

#include <hdf5.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return QL_ERR_ARG;
}

// General_Definitions.h:236: PI = atan(1)*4, which rounds to this double
constexpr double QL_PI = 3.14159265358979323846;

// closes an HDF5 handle on scope exit
struct H5Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H5Handle()
    {
        if (id >= 0) close(id);
    }
    H5Handle(const H5Handle &) = delete;
    H5Handle &operator=(const H5Handle &) = delete;
};

bool check_layout(const qlio_layout *L)
{
    return L && L->NX >= 2 && L->NY >= 2 && L->NZ >= 4 && L->nxy == (L->NX - 1) * (L->NY / 2) &&
           L->nz == 1 + 2 * (int)(L->NZ / 3.0) && L->nlin > 0 && L->nmf > 0 && L->L[0] > 0 && L->L[1] > 0 &&
           L->L[2] > 0;
}

// mode index tables and wavenumbers (Auxiliary/Kdata.cpp:28-80)
struct Tables {
    std::vector<int> kxi, kyi;
    std::vector<double> ky, kz;
};
Tables tables(const qlio_layout &L)
{
    Tables T;
    const int nx = L.NX / 2, ny = L.NY / 2;
    T.kxi.assign(L.nxy, 0);
    T.kyi.assign(L.nxy, 0);
    T.ky.assign(L.nxy, 0.0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            T.kxi[j + ny * i] = i;
            T.kyi[j + ny * i] = j;
        }
    for (int i = nx + 1; i < 2 * nx; ++i)
        for (int j = 0; j < ny; ++j) {
            T.kxi[j + ny * (i - 1)] = -2 * nx + i;
            T.kyi[j + ny * (i - 1)] = j;
        }
    const double kyL = 2 * QL_PI / L.L[1], kzL = 2 * QL_PI / L.L[2];
    for (int k = 0; k < L.nxy; ++k) T.ky[k] = T.kyi[k] * kyL;
    const int nzl = (L.nz - 1) / 2;
    T.kz.assign(L.nz, 0.0);
    for (int i = 0; i < nzl + 1; ++i) T.kz[i] = i * kzL;
    for (int i = nzl + 1; i < 2 * nzl + 1; ++i) T.kz[i] = (-2 * nzl + i - 1) * kzL;
    return T;
}

std::string mode_name(const Tables &T, int jj)
{
    char b[96];
    snprintf(b, sizeof(b), "Ind %d: kx %d ky %d", jj, T.kxi[jj], T.kyi[jj]);
    return b;
}

int write_dataset(hid_t loc, const char *name, int rank, const hsize_t *dims, const double *data, bool fp64)
{
    H5Handle space(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    if (space.id < 0) return fail("H5Screate_simple failed for %s", name);
    H5Handle ds(H5Dcreate2(loc, name, fp64 ? H5T_IEEE_F64LE : H5T_IEEE_F32LE, space.id, H5P_DEFAULT, H5P_DEFAULT,
                           H5P_DEFAULT),
                H5Dclose);
    if (ds.id < 0) return fail("cannot create dataset %s", name);
    herr_t e;
    if (fp64) {
        e = H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    } else {
        size_t n = 1;
        for (int r = 0; r < rank; ++r) n *= dims[r];
        std::vector<float> f(n);
        for (size_t k = 0; k < n; ++k) f[k] = (float)data[k];
        e = H5Dwrite(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, f.data());
    }
    if (e < 0) return fail("cannot write dataset %s", name);
    return QL_OK;
}

int read_dataset(hid_t loc, const char *name, size_t n, double *out)
{
    H5Handle ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) return fail("dataset %s not found", name);
    H5Handle space(H5Dget_space(ds.id), H5Sclose);
    const hssize_t np = H5Sget_simple_extent_npoints(space.id);
    if (np < 0 || (size_t)np != n) return fail("dataset %s has %lld values, expected %zu", name, (long long)np, n);
    if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail("cannot read dataset %s", name);
    return QL_OK;
}

// names of the "t=..." groups in the root group with their times
struct TimeGroups {
    std::vector<std::string> names;
    std::vector<double> t;
};
herr_t collect_time_group(hid_t loc, const char *name, const H5L_info_t *, void *op)
{
    H5O_info_t oi;
    if (H5Oget_info_by_name(loc, name, &oi, H5P_DEFAULT) < 0 || oi.type != H5O_TYPE_GROUP) return 0;
    // FullSave_Load.cpp:401-415: group names are "t=" followed by the time
    if (strncmp(name, "t=", 2) != 0) return 0;
    auto *tg = static_cast<TimeGroups *>(op);
    tg->names.push_back(name);
    tg->t.push_back(strtod(name + 2, nullptr));
    return 0;
}

int time_groups(hid_t f, TimeGroups *tg)
{
    hsize_t idx = 0;
    if (H5Literate(f, H5_INDEX_NAME, H5_ITER_NATIVE, &idx, collect_time_group, tg) < 0)
        return fail("cannot list the groups of the file");
    return QL_OK;
}

std::string join(const char *dir, const char *file)
{
    std::string d(dir ? dir : "");
    if (!d.empty() && d.back() != '/') d += '/';
    return d + file;
}

const char *const DAT_FILES[] = {"energy.dat", "angular_momentum.dat", "dissipation.dat", "reynolds_stress.dat",
                                 "mean_fields.dat", "time_vec.dat"};

}  // namespace

extern "C" {

const char *qlio_last_error(void) { return g_err.c_str(); }

int qlio_layout_from_params(const ql_params *p, qlio_layout *out)
{
    if (!p || !out) return fail("qlio_layout_from_params: null argument");
    qlio_layout L{};
    L.NX = p->NX;
    L.NY = p->NY;
    L.NZ = p->NZ;
    L.nxy = (p->NX - 1) * (p->NY / 2);
    L.nz = 1 + 2 * (int)(p->NZ / 3.0);
    L.nlin = (p->model == QL_MODEL_HD_FULLU) ? 2 : 4;
    L.nmf = (p->model == QL_MODEL_MHD_FULLQLIN) ? 4 : 2;
    for (int k = 0; k < 3; ++k) L.L[k] = p->L[k];
    if (!check_layout(&L)) return fail("qlio_layout_from_params: invalid sizes");
    *out = L;
    return QL_OK;
}

int qlio_h5_save(const char *path, int create, const qlio_layout *L, double t, double q, const double *lin,
                 const double *mf, const double *kx0, int fp64)
{
    if (!path || !lin || !mf || !kx0) return fail("qlio_h5_save: null argument");
    if (!check_layout(L)) return fail("qlio_h5_save: invalid layout");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    hid_t fid = -1;
    if (!create) {
        fid = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
        if (fid < 0) create = 1;
    }
    if (create) fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fid < 0) return fail("qlio_h5_save: cannot open or create %s", path);
    H5Handle f(fid, H5Fclose);
    const Tables T = tables(*L);
    int rc;
    if (create) {
        // the time-independent tables (FullSave_Load.cpp:114-148)
        const hsize_t dxy = (hsize_t)L->nxy, dz = (hsize_t)L->nz;
        if ((rc = write_dataset(f.id, "ky", 1, &dxy, T.ky.data(), false))) return rc;
        if ((rc = write_dataset(f.id, "kz", 1, &dz, T.kz.data(), true))) return rc;
    }
    char gname[64];
    snprintf(gname, sizeof(gname), "/t=%.15g", t);
    H5Handle g(H5Gcreate2(f.id, gname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (g.id < 0) return fail("qlio_h5_save: cannot create group %s (already saved?)", gname);
    const hsize_t dlin[2] = {(hsize_t)L->nz * L->nlin, 2}, dmf[2] = {(hsize_t)L->nz * L->nmf, 2};
    const size_t per = (size_t)2 * L->nlin * L->nz;
    for (int jj = 0; jj < L->nxy; ++jj)
        if ((rc = write_dataset(g.id, mode_name(T, jj).c_str(), 2, dlin, lin + jj * per, fp64 != 0))) return rc;
    if ((rc = write_dataset(g.id, "Mean", 2, dmf, mf, fp64 != 0))) return rc;
    std::vector<double> kx(L->nxy);
    for (int j = 0; j < L->nxy; ++j) kx[j] = kx0[j] + q * t * T.ky[j];  // FullSave_Load.cpp:231-232
    const hsize_t dxy = (hsize_t)L->nxy;
    if ((rc = write_dataset(g.id, "kx", 1, &dxy, kx.data(), fp64 != 0))) return rc;
    if (fp64) {
        if ((rc = write_dataset(g.id, "kx0", 1, &dxy, kx0, true))) return rc;
        H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
        H5Handle at(H5Acreate2(g.id, "t", H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        if (at.id < 0 || H5Awrite(at.id, H5T_NATIVE_DOUBLE, &t) < 0) return fail("qlio_h5_save: cannot write t");
    }
    return QL_OK;
}

int qlio_h5_load_latest(const char *path, const qlio_layout *L, double q, double *t, double *lin, double *mf,
                        double *kx0)
{
    if (!path) return fail("qlio_h5_load_latest: null path");
    if (!check_layout(L)) return fail("qlio_h5_load_latest: invalid layout");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return fail("qlio_h5_load_latest: cannot open %s", path);
    TimeGroups tg;
    int rc;
    if ((rc = time_groups(f.id, &tg))) return rc;
    if (tg.names.empty()) return fail("qlio_h5_load_latest: no time save in %s", path);
    size_t best = 0;
    for (size_t k = 1; k < tg.t.size(); ++k)
        if (tg.t[k] > tg.t[best]) best = k;
    H5Handle g(H5Gopen2(f.id, tg.names[best].c_str(), H5P_DEFAULT), H5Gclose);
    if (g.id < 0) return fail("qlio_h5_load_latest: cannot open group %s", tg.names[best].c_str());
    double tt = tg.t[best];
    if (H5Aexists(g.id, "t") > 0) {
        H5Handle at(H5Aopen(g.id, "t", H5P_DEFAULT), H5Aclose);
        if (at.id < 0 || H5Aread(at.id, H5T_NATIVE_DOUBLE, &tt) < 0) return fail("qlio_h5_load_latest: bad t");
    }
    const Tables T = tables(*L);
    if (lin) {
        const size_t per = (size_t)2 * L->nlin * L->nz;
        for (int jj = 0; jj < L->nxy; ++jj)
            if ((rc = read_dataset(g.id, mode_name(T, jj).c_str(), per, lin + jj * per))) return rc;
    }
    if (mf && (rc = read_dataset(g.id, "Mean", (size_t)2 * L->nmf * L->nz, mf))) return rc;
    if (kx0) {
        if (H5Lexists(g.id, "kx0", H5P_DEFAULT) > 0) {
            if ((rc = read_dataset(g.id, "kx0", (size_t)L->nxy, kx0))) return rc;
        } else {
            // FullSave_Load.cpp:325-333: kx = saved kx - q t_start ky
            if ((rc = read_dataset(g.id, "kx", (size_t)L->nxy, kx0))) return rc;
            for (int j = 0; j < L->nxy; ++j) kx0[j] = kx0[j] - q * tt * T.ky[j];
        }
    }
    if (t) *t = tt;
    return QL_OK;
}

int qlio_h5_count(const char *path, int *ngroups)
{
    if (!path || !ngroups) return fail("qlio_h5_count: null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H5Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return fail("qlio_h5_count: cannot open %s", path);
    TimeGroups tg;
    int rc;
    if ((rc = time_groups(f.id, &tg))) return rc;
    *ngroups = (int)tg.names.size();
    return QL_OK;
}

int qlio_dat_init(const char *dir)
{
    if (!dir) return fail("qlio_dat_init: null dir");
    for (const char *name : DAT_FILES) {
        std::ofstream o(join(dir, name), std::ios::binary | std::ios::trunc);
        if (!o) return fail("qlio_dat_init: cannot create %s", join(dir, name).c_str());
    }
    return QL_OK;
}

int qlio_dat_append_values(const char *dir, double t, const double energy[4], const double am[4],
                           const double diss[4], const double *reynolds, int nrey, const double *mf_real, int nmf,
                           int NZ)
{
    if (!dir || !energy || !am || !diss || (nrey > 0 && !reynolds)) return fail("qlio_dat_append: null argument");
    if (nrey < 0 || (mf_real && (nmf < 1 || NZ < 1))) return fail("qlio_dat_append: bad sizes");
    auto app = [&](const char *name, const double *v, size_t n) {
        std::ofstream o(join(dir, name), std::ios::binary | std::ios::app);
        if (!o) return fail("qlio_dat_append: cannot open %s", join(dir, name).c_str());
        o.write(reinterpret_cast<const char *>(v), (std::streamsize)(sizeof(double) * n));
        return o ? QL_OK : fail("qlio_dat_append: write to %s failed", join(dir, name).c_str());
    };
    int rc;
    if ((rc = app("energy.dat", energy, 4))) return rc;
    if ((rc = app("angular_momentum.dat", am, 4))) return rc;
    if ((rc = app("dissipation.dat", diss, 4))) return rc;
    if ((rc = app("reynolds_stress.dat", reynolds, (size_t)nrey))) return rc;
    if ((rc = app("time_vec.dat", &t, 1))) return rc;
    if (mf_real && (rc = app("mean_fields.dat", mf_real, (size_t)nmf * NZ))) return rc;
    return QL_OK;
}

int qlio_dat_append(const char *dir, double t, const double diag[18], int nrey, const double *mf_real, int nmf,
                    int NZ)
{
    if (!dir || !diag) return fail("qlio_dat_append: null argument");
    if (nrey < 0) return fail("qlio_dat_append: bad sizes");
    std::vector<double> rey(nrey, 0.0);
    for (int k = 0; k < nrey && k < 4; ++k) rey[k] = diag[12 + k];
    return qlio_dat_append_values(dir, t, diag, diag + 4, diag + 8, rey.data(), nrey, mf_real, nmf, NZ);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// reference-shaped classes over the GPU model
// ---------------------------------------------------------------------------
namespace qlgpu {

namespace {
qlio_layout layout_of(GpuModel *eqs)
{
    const Inputs &sp = eqs->inputs();
    qlio_layout L{};
    L.NX = sp.NXY[0];
    L.NY = sp.NXY[1];
    L.NZ = sp.NZ;
    L.nxy = eqs->Dimxy();
    L.nz = eqs->NZ();
    L.nlin = eqs->num_Lin();
    L.nmf = eqs->num_MFs();
    for (int k = 0; k < 3; ++k) L.L[k] = sp.L[k];
    return L;
}
void io_check(int rc)
{
    if (rc != QL_OK) throw Error(rc, qlio_last_error());
}
}  // namespace

FullSave_Load::FullSave_Load(const std::string &simulation_dir, bool saveQ, bool start_from_savedQ, GpuModel *eqs,
                             bool fp64)
    : saveQ_(saveQ), start_from_saved_Q_(start_from_savedQ), fp64_(fp64), eqs_(eqs),
      file_name_(simulation_dir + "FullSolution.h5")
{
    // FullSave_Load.cpp:24-66: restart needs an existing file; otherwise a new one
    if (start_from_saved_Q_) {
        int n = 0;
        if (qlio_h5_count(file_name_.c_str(), &n) != QL_OK || n == 0) start_from_saved_Q_ = false;
    }
    fresh_ = !start_from_saved_Q_;
}

void FullSave_Load::SaveSolution(double t, solution *sol)
{
    if (!saveQ_) return;
    const qlio_layout L = layout_of(eqs_);
    std::vector<dcmplx> lin, mf;
    sol->pack(lin, mf);
    eqs_->push_kx();
    const std::vector<double> kx0 = eqs_->kx();
    io_check(qlio_h5_save(file_name_.c_str(), fresh_ ? 1 : 0, &L, t, eqs_->inputs().q,
                          reinterpret_cast<const double *>(lin.data()), reinterpret_cast<const double *>(mf.data()),
                          kx0.data(), fp64_ ? 1 : 0));
    fresh_ = false;
}

double FullSave_Load::LoadSolution_forRestart(solution *sol)
{
    const qlio_layout L = layout_of(eqs_);
    std::vector<dcmplx> lin((size_t)L.nxy * L.nlin * L.nz), mf((size_t)L.nmf * L.nz);
    std::vector<double> kx0(L.nxy);
    double t = 0.0;
    io_check(qlio_h5_load_latest(file_name_.c_str(), &L, eqs_->inputs().q, &t, reinterpret_cast<double *>(lin.data()),
                                 reinterpret_cast<double *>(mf.data()), kx0.data()));
    sol->unpack(lin, mf);
    eqs_->set_kx(kx0.data());  // K->kx and the device (FullSave_Load.cpp:344)
    return t;
}

TimeVariablesFiles::TimeVariablesFiles(const std::string &simulation_dir, bool append) : dir_(simulation_dir)
{
    // TimeVariables.cpp:67-118: truncate unless restarting
    if (!append) io_check(qlio_dat_init(dir_.c_str()));
}

void TimeVariablesFiles::Save_Data(double t, TimeVariables *tv, int nrey)
{
    io_check(qlio_dat_append_values(dir_.c_str(), t, tv->current_energy(), tv->current_AM(), tv->current_diss(),
                                    tv->current_reynolds(nrey), nrey, nullptr, 0, 0));
}

void TimeVariablesFiles::Save_Mean_Fields(GpuModel *eqs)
{
    std::vector<double> out((size_t)eqs->num_MFs() * eqs->NZfull());
    eqs->check(ql_mf_real(eqs->ctx(), out.data()));
    std::ofstream o(dir_ + "mean_fields.dat", std::ios::binary | std::ios::app);
    o.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)(sizeof(double) * out.size()));
    if (!o) throw Error(QL_ERR_ARG, "cannot append to " + dir_ + "mean_fields.dat");
}

}  // namespace qlgpu
