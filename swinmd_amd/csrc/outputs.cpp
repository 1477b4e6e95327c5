// outputs.cpp -- the reference's restart and pressure output files (host
// code, no GPU), written with the same stream formatting the reference uses.
//
//   state file     Integrator::write_state            src/Integrator.cpp:275-302
//   vtMinus file   NHIntegrator::write_vtMinus        src/integrators/NHIntegrator.cpp:1782-1803
//   pressure lines Ensemble::write_press_res / write_pressac / write_pressureTensor
//                  src/Ensemble.cpp:1153-1172, :1887-1912
//
// Numbers go through std::ostream with the default float format (precision 6
// unless a caller asks for more, e.g. 17 for a lossless restart), std::setw
// and std::scientific where the reference sets them.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>

#include "../../include/swinmd_hip.h"

namespace {

void put_err(char* err, size_t errlen, const std::string& m) {
    if (err && errlen) {
        std::strncpy(err, m.c_str(), errlen - 1);
        err[errlen - 1] = '\0';
    }
}

template <class F>
int write_file(const char* path, bool append, char* err, size_t errlen, F&& body) {
    if (!path) return SMD_EINVAL;
    std::ofstream of(path, append ? std::ios::app : std::ios::out);
    if (!of.is_open()) {
        put_err(err, errlen, std::string("can't open ") + path + " for writing");
        return SMD_EIO;
    }
    body(of);
    of.close();
    if (of.fail()) {
        put_err(err, errlen, std::string("write to ") + path + " failed");
        return SMD_EIO;
    }
    put_err(err, errlen, "");
    return SMD_OK;
}

void block(std::ofstream& of, int64_t n, const double* v) {
    of << n << std::endl;
    for (int64_t j = 0; j < n; ++j) of << v[3 * j] << ' ' << v[3 * j + 1] << ' ' << v[3 * j + 2] << std::endl;
}

enum { XX = 0, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ };

}  // namespace

extern "C" int smd_write_state(const char* path, int64_t step, int64_t n_atoms, const double* pos,
                               const double* real_pos, const double* vel, int32_t precision, char* err,
                               size_t errlen) {
    if (!pos || !real_pos || !vel || n_atoms < 0 || precision < 0) return SMD_EINVAL;
    return write_file(path, false, err, errlen, [&](std::ofstream& of) {
        if (precision) of << std::setprecision(precision);
        of << "Coordinate at step: " << step << std::endl;
        block(of, n_atoms, pos);
        of << "Real positions at step: " << step << std::endl;
        block(of, n_atoms, real_pos);
        of << "velocity at steps: " << step << std::endl;
        block(of, n_atoms, vel);
        of << std::endl;
    });
}

extern "C" int smd_write_vt_minus(const char* path, int64_t step, int64_t n_atoms, const double* v, int32_t precision,
                                  char* err, size_t errlen) {
    if (!v || n_atoms < 0 || precision < 0) return SMD_EINVAL;
    return write_file(path, false, err, errlen, [&](std::ofstream& of) {
        if (precision) of << std::setprecision(precision);
        of << "vtMinus" << step << std::endl;
        block(of, n_atoms, v);
        of << std::endl;
    });
}

extern "C" int smd_write_press_res(const char* path, int append, int64_t step, const double* ap, const double* mp,
                                   char* err, size_t errlen) {
    if (!ap || !mp) return SMD_EINVAL;
    using std::setw;
    return write_file(path, append != 0, err, errlen, [&](std::ofstream& of) {
        of << step << setw(14) << (ap[XX] + ap[YY] + ap[ZZ]) / 3.0 << setw(14) << (ap[XY] + ap[YX]) * 0.5;
        of << setw(14) << (ap[YX] - ap[XY]) * 0.5 << setw(14) << (ap[XX] - ap[YY]) << setw(14) << (ap[YY] - ap[ZZ]);
        of << setw(14) << (mp[XX] + mp[YY] + mp[ZZ]) / 3.0 << setw(14) << (mp[XY] + mp[YX]) * 0.5;
        of << setw(14) << (mp[YX] - mp[XY]) * 0.5 << setw(14) << (mp[XX] - mp[YY]) << setw(14) << (mp[YY] - mp[ZZ]);
        of << std::endl;
    });
}

extern "C" int smd_write_pressac(const char* path, int append, int64_t step, const double* atom_stress,
                                 const double* mol_stress, int64_t samples, char* err, size_t errlen) {
    if (!atom_stress || !mol_stress) return SMD_EINVAL;
    using std::setw;
    return write_file(path, append != 0, err, errlen, [&](std::ofstream& of) {
        if (samples <= 0) return;   // luscount > 0 (Ensemble.cpp:1163)
        const double c = (double)samples;
        of << step;
        for (int k = 0; k < 5; ++k) of << setw(14) << atom_stress[k] / c;
        for (int k = 0; k < 5; ++k) of << setw(14) << mol_stress[k] / c;
        of << std::endl;
    });
}

extern "C" int smd_write_pressure_tensor(const char* path, int append, int64_t counter1, const double* ap,
                                         const double* mp, double atom_shear_sum, double mol_shear_sum,
                                         int64_t counter, char* err, size_t errlen) {
    if (!ap || !mp || counter == 0) return SMD_EINVAL;
    using std::setw;
    return write_file(path, append != 0, err, errlen, [&](std::ofstream& of) {
        of << counter1 << '\t' << std::endl;
        auto tensor = [&](const double* p, const char* pre) {
            of << std::scientific << "  " << pre << "XX = " << p[XX] << "  " << pre << "YY = " << p[YY] << "  " << pre
               << "ZZ = " << p[ZZ];
            of << std::scientific << "  " << pre << "XY = " << p[XY] << "  " << pre << "YX = " << p[YX] << "  " << pre
               << "XZ = " << p[XZ] << "  " << pre << "ZX = " << p[ZX];
            of << std::scientific << "  " << pre << "YZ = " << p[YZ] << "  " << pre << "ZY = " << p[ZY];
            of << std::scientific << setw(14) << (p[XY] + p[YX]) * 0.5;
            of << std::scientific << setw(14) << (p[ZY] + p[YZ]) * 0.5;
            of << std::scientific << setw(14) << (p[XZ] + p[ZX]) * 0.5;
            of << std::scientific << setw(14) << (p[XZ] + p[ZX] + p[ZY] + p[YZ] + p[XY] + p[YX]) / 6.0 << std::endl;
        };
        tensor(ap, "");
        tensor(mp, "mol");
        of << std::scientific << setw(14) << atom_shear_sum / (double)counter;
        of << std::scientific << setw(14) << mol_shear_sum / (double)counter << std::endl;
    });
}
