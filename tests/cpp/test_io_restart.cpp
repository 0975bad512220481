// Dump and restart through the reference-shaped file classes
// (FullSave_Load, TimeVariables' .dat files) the way main.cpp uses them
// (main.cpp:79-160): run A steps n1, saves the full solution, steps n2 more;
// run B starts from the save (LoadSolution_forRestart, a new integrator at
// t_start) and steps n2.  Box 16x16x32, L = (1, 4, 1), q = 1.5, Omega = 1,
// nu = eta = 1/2000, dt = 0.01, RK3CN_Elin, remap on.
//
// usage: test_io_restart <dir/> <state.bin> <n1> <n2> <fp64> <out.bin>
// out.bin: for A then B: lin, mf (complex128), kx (double), t.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../ql_dns_amd/csrc/io/ql_io.hpp"

using namespace qlgpu;

namespace {
Inputs box()
{
    Inputs SP;
    SP.NXY[0] = 16;
    SP.NXY[1] = 16;
    SP.NZ = 32;
    SP.L[0] = 1.0;
    SP.L[1] = 4.0;
    SP.L[2] = 1.0;
    SP.nu = SP.eta = 1.0 / 2000;
    SP.dt = 0.01;
    return SP;
}

void dump(std::FILE *fo, GpuModel *m, solution *sol, double t)
{
    const int nxy = m->Dimxy(), nz = m->NZ();
    for (int i = 0; i < nxy; ++i)
        for (int j = 0; j < 4; ++j) std::fwrite(sol->pLin(i, j)->data(), sizeof(dcmplx), nz, fo);
    for (int f = 0; f < m->num_MFs(); ++f) std::fwrite(sol->pMF(f)->data(), sizeof(dcmplx), nz, fo);
    std::vector<double> kx = m->kx();
    std::fwrite(kx.data(), sizeof(double), nxy, fo);
    std::fwrite(&t, sizeof(double), 1, fo);
}

// main.cpp:108-123 plus the time-variable output at every step
double run(GpuModel *m, solution *sol, Integrator *integ, TimeVariables *tv, TimeVariablesFiles *files, double t,
           int n)
{
    const Inputs &SP = m->inputs();
    for (int k = 0; k < n; ++k) {
        const double dt = integ->Step(t, sol);
        t += dt;
        if (SP.remapQ) m->ShearingBox_Remap(SP.q * t, sol);
        if (files) {
            m->Calc_Energy_AM_Diss(tv, t, sol);
            files->Save_Data(t, tv, m->num_reynolds_saves());
            files->Save_Mean_Fields(m);
        }
    }
    return t;
}
}  // namespace

int main(int argc, char **argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s dir/ state.bin n1 n2 fp64 out.bin\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int n1 = std::atoi(argv[3]), n2 = std::atoi(argv[4]);
    const bool fp64 = std::atoi(argv[5]) != 0;
    try {
        const Inputs SP = box();
        std::FILE *fo = std::fopen(argv[6], "wb");
        {
            GpuModel m(SP);
            solution sol(&m);
            std::FILE *fi = std::fopen(argv[2], "rb");
            if (!fi) return 3;
            for (int i = 0; i < m.Dimxy(); ++i)
                for (int j = 0; j < 4; ++j)
                    if (std::fread(sol.pLin(i, j)->data(), sizeof(dcmplx), m.NZ(), fi) != (size_t)m.NZ()) return 3;
            for (int f = 0; f < m.num_MFs(); ++f)
                if (std::fread(sol.pMF(f)->data(), sizeof(dcmplx), m.NZ(), fi) != (size_t)m.NZ()) return 3;
            std::fclose(fi);
            TimeVariables tv;
            TimeVariablesFiles files(dir, false);
            FullSave_Load save(dir, true, false, &m, fp64);
            double t = SP.t_start;
            GpuIntegrator integ(t, SP, &m);
            m.Calc_Energy_AM_Diss(&tv, t, &sol);
            files.Save_Data(t, &tv, m.num_reynolds_saves());
            files.Save_Mean_Fields(&m);
            t = run(&m, &sol, &integ, &tv, &files, t, n1);
            save.SaveSolution(t, &sol);
            t = run(&m, &sol, &integ, &tv, &files, t, n2);
            dump(fo, &m, &sol, t);
        }
        {
            GpuModel m(SP);
            solution sol(&m);
            FullSave_Load load(dir, false, true, &m, fp64);
            if (!load.start_from_saved()) return 5;
            double t = load.LoadSolution_forRestart(&sol);
            GpuIntegrator integ(t, SP, &m);
            t = run(&m, &sol, &integ, nullptr, nullptr, t, n2);
            dump(fo, &m, &sol, t);
        }
        std::fclose(fo);
    } catch (const Error &e) {
        std::fprintf(stderr, "qlgpu::Error %d: %s\n", e.code, e.what());
        return 4;
    }
    return 0;
}
