// ql_io.hpp -- reference-shaped file output over the GPU model (libql_io.so):
// FullSave_Load (Auxiliary/FullSave_Load.h:22-80) and the file side of
// TimeVariables (Auxiliary/TimeVariables.cpp:52-228).
#pragma once
#include <string>

#include "../../../include/ql_io.h"
#include "../host/ql_model.hpp"

namespace qlgpu {

// HDF5 dump and restart of the full solution, FullSolution.h5 in the
// simulation directory (the directory string ends with '/', as the
// reference's simulation_dir does).  fp64 writes the float64 variant.
class FullSave_Load {
public:
    FullSave_Load(const std::string &simulation_dir, bool saveQ, bool start_from_savedQ, GpuModel *eqs,
                  bool fp64 = false);
    // FullSave_Load.cpp:183-252
    void SaveSolution(double t, solution *sol);
    // FullSave_Load.cpp:264-347: loads the latest save into sol and the
    // model's kx; returns the time of that save (the reference's t_start)
    double LoadSolution_forRestart(solution *sol);
    bool start_from_saved() const { return start_from_saved_Q_; }

private:
    bool saveQ_, start_from_saved_Q_, fp64_, fresh_;
    GpuModel *eqs_;
    std::string file_name_;
};

// The .dat time series of TimeVariables, appended per save.
class TimeVariablesFiles {
public:
    TimeVariablesFiles(const std::string &simulation_dir, bool append);
    // TimeVariables::Save_Data: energy, angular momentum, dissipation, the
    // model's num_reynolds_saves() Reynolds values, time
    void Save_Data(double t, TimeVariables *tv, int nrey);
    // TimeVariables::Save_Mean_Fields: fac1D Re(ifft(pad(MF))) per mean field
    void Save_Mean_Fields(GpuModel *eqs);

private:
    std::string dir_;
};

}  // namespace qlgpu
