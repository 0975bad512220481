// Drives the GPU path through the C++ Model / Integrator / solution mirror the
// way the reference's main.cpp does (main.cpp:79-123: solution, initial state,
// integrator construction, Calc_Energy_AM_Diss, the Step / t += dt / remap
// loop), plus one Model::rhs call.
//
// usage: test_host_adapter <model> <integrator> <nsteps> <state.bin> <out.bin> [dt CFL]
// box: 16x16x32, L = (1, 4, 1), q = 1.5, Omega = 1, nu = eta = 1/2000, dt = 0.01
// unless given (dt < 0: variable steps of at most -dt, RK3CN.cpp:20-30).
// state.bin: lin [nxy][4][nz] then mf [nMF][nz], complex128.
// out.bin: rhs_lin, rhs_mf, linop (from the initial state at t = 0.0, dt_lin = 0.0053),
//          then lin, mf, kx, t, 18 diagnostics after nsteps, the Reynolds-save
//          count and values, then the Kdata tables Model::K holds: Im(ky),
//          kx_index_full, ky_index_full (per mode), Im(kz), kz2 (per kz slot),
//          lap2 and ilap2 of every ky index.  Before the loop the last mode's
//          K->kx is moved by half a kx spacing through Model::K (as a restart
//          writes it, FullSave_Load.cpp:344); kx is read back from K->kx.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../ql_dns_amd/csrc/host/ql_model.hpp"

using namespace qlgpu;

int main(int argc, char **argv)
{
    if (argc != 6 && argc != 8) {
        std::fprintf(stderr, "usage: %s model integrator nsteps state.bin out.bin [dt CFL]\n", argv[0]);
        return 2;
    }
    try {
        Inputs SP;
        SP.equations_to_use = argv[1];
        SP.integrator = argv[2];
        const int nsteps = std::atoi(argv[3]);
        SP.NXY[0] = 16;
        SP.NXY[1] = 16;
        SP.NZ = 32;
        SP.L[0] = 1.0;
        SP.L[1] = 4.0;
        SP.L[2] = 1.0;
        SP.nu = SP.eta = 1.0 / 2000;
        SP.dt = argc == 8 ? std::atof(argv[6]) : 0.01;
        SP.CFL = argc == 8 ? std::atof(argv[7]) : 1.0;
        SP.q = 1.5;
        SP.omega = 1.0;

        GpuModel *fluidEqs = new GpuModel(SP);
        const int nxy = fluidEqs->Dimxy(), nz = fluidEqs->NZ(), nMF = fluidEqs->num_MFs(), nLin = fluidEqs->num_Lin();
        solution *sol = new solution(fluidEqs);
        solution *sol_rhs = new solution(fluidEqs);

        std::FILE *fi = std::fopen(argv[4], "rb");
        if (!fi) return 3;
        for (int i = 0; i < nxy; ++i)
            for (int j = 0; j < nLin; ++j)
                if (std::fread(sol->pLin(i, j)->data(), sizeof(dcmplx), nz, fi) != (size_t)nz) return 3;
        for (int f = 0; f < nMF; ++f)
            if (std::fread(sol->pMF(f)->data(), sizeof(dcmplx), nz, fi) != (size_t)nz) return 3;
        std::fclose(fi);

        std::FILE *fo = std::fopen(argv[5], "wb");
        // Model::rhs with caller-allocated linear-operator arrays (RK3CN.cpp:41-50)
        std::vector<doubVec> Lstore((size_t)nxy * nLin, doubVec(nz));
        std::vector<doubVec *> Lrows(nxy);
        for (int i = 0; i < nxy; ++i) Lrows[i] = &Lstore[(size_t)i * nLin];
        fluidEqs->rhs(0.0, 0.0053, sol, sol_rhs, Lrows.data());
        for (int i = 0; i < nxy; ++i)
            for (int j = 0; j < nLin; ++j) std::fwrite(sol_rhs->pLin(i, j)->data(), sizeof(dcmplx), nz, fo);
        for (int f = 0; f < nMF; ++f) std::fwrite(sol_rhs->pMF(f)->data(), sizeof(dcmplx), nz, fo);
        for (int i = 0; i < nxy; ++i)
            for (int j = 0; j < nLin; ++j) std::fwrite(Lrows[i][j].data(), sizeof(double), nz, fo);

        // a write through Model::K, as FullSave_Load's restart does
        fluidEqs->K->kx[nxy - 1] = dcmplx(0, fluidEqs->K->kx[nxy - 1].imag() + 0.5 * fluidEqs->K->kxLfac);

        // main.cpp:90-123
        double t = SP.t_start, dt;
        Integrator *integrator = new GpuIntegrator(t, SP, fluidEqs);
        TimeVariables tv;
        fluidEqs->Calc_Energy_AM_Diss(&tv, t, sol);
        for (int n = 0; n < nsteps; ++n) {
            dt = integrator->Step(t, sol);
            if (SP.f_noise != 0.0) fluidEqs->DrivingNoise(t, dt, sol);
            t += dt;
            if (SP.remapQ) fluidEqs->ShearingBox_Remap(SP.q * t, sol);
        }
        fluidEqs->Calc_Energy_AM_Diss(&tv, t, sol);
        sol->Check_Solution();  // main.cpp:136; throws where the reference aborts

        for (int i = 0; i < nxy; ++i)
            for (int j = 0; j < nLin; ++j) std::fwrite(sol->pLin(i, j)->data(), sizeof(dcmplx), nz, fo);
        for (int f = 0; f < nMF; ++f) std::fwrite(sol->pMF(f)->data(), sizeof(dcmplx), nz, fo);
        std::vector<double> kx(nxy);
        for (int i = 0; i < nxy; ++i) kx[i] = fluidEqs->K->kx[i].imag();
        std::fwrite(kx.data(), sizeof(double), nxy, fo);
        std::fwrite(&t, sizeof(double), 1, fo);
        const TimeVariables::Record &r = tv.records().back();
        double d[18];
        for (int k = 0; k < 4; ++k) {
            d[k] = r.energy[k];
            d[4 + k] = r.AM[k];
            d[8 + k] = r.diss[k];
            d[12 + k] = r.reynolds[k];
        }
        d[16] = tv.etaK_times_kmax[0];
        d[17] = tv.etaK_times_kmax[1];
        std::fwrite(d, sizeof(double), 18, fo);
        // the model's num_reynolds_saves values (HD_fullU: 2 NZ stress profile values)
        const double nrey = (double)fluidEqs->num_reynolds_saves();
        std::fwrite(&nrey, sizeof(double), 1, fo);
        std::fwrite(r.reynolds.data(), sizeof(double), (size_t)nrey, fo);
        const Kdata *K = fluidEqs->K;
        for (int i = 0; i < nxy; ++i) {
            const double v[3] = {K->ky[i].imag(), (double)K->kx_index_full[i], (double)K->ky_index_full[i]};
            std::fwrite(v, sizeof(double), 3, fo);
        }
        for (int m = 0; m < nz; ++m) {
            const double v[2] = {K->kz(m).imag(), K->kz2(m)};
            std::fwrite(v, sizeof(double), 2, fo);
        }
        for (int j = 0; j < SP.NXY[1] / 2; ++j) {
            std::fwrite(K->lap2[j].data(), sizeof(double), nz, fo);
            std::fwrite(K->ilap2[j].data(), sizeof(double), nz, fo);
        }
        std::fclose(fo);
        std::printf("mean time step %.6g over %d steps\n", integrator->mean_time_step(), nsteps);
        delete integrator;
        delete sol_rhs;
        delete sol;
        delete fluidEqs;
    } catch (const Error &e) {
        std::fprintf(stderr, "qlgpu::Error %d: %s\n", e.code, e.what());
        return 4;
    }
    return 0;
}
