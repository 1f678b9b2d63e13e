// The reference's HDF5 formats (include/gpmood_io.h): host code over the HDF5
// C library.  Built as libgpmood_io.so, apart from the step library.
#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpmood_io.h"

namespace {

constexpr double kGamma = 1.4;  // constants.f90:9

// Names of compute_metadata (output.f90:13-72); empty when the reference has none.
std::string problem_name(int id)
{
    switch (id) {
    case 28: return "2DRP3";
    case 49: return "2DRP4";
    case 38: return "2DRP6";
    case 30: return "2DRP12";
    case 31: return "2DRP15";
    case 50: return "2DRP17";
    case 29: return "implosion";
    case 39: return "sedov";
    case 44: return "shu_osher";
    case 35: return "DMR";
    case 36: return "explosion";
    case 45: return "Mach800";
    case 51: return "RT";
    case 41: return "KH";
    default: return "";
    }
}

std::string trimmed(const char *s)
{
    std::string r = s ? s : "";
    const size_t a = r.find_first_not_of(' '), b = r.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : r.substr(a, b - a + 1);
}

std::string method_name(const gpm_io_meta &m)
{
    switch (m.method) {
    case 2: return m.dmp ? "GP_MOOD" : "GP_MOOD_noDMP";
    case 200: return "POL_MOOD";
    case 203: return "NN_GP_MOOD_" + trimmed(m.nn_filename);
    case 204: return "NN_GP_MOOD_CC_" + trimmed(m.nn_filename);
    case 1: return "FOG";
    default: return "";
    }
}

// write(CFL_char, "(F3.1)") CFL: three characters, one decimal; asterisks
// when the value does not fit.
std::string cfl_name(double cfl)
{
    char b[64];
    std::snprintf(b, sizeof(b), "%3.1f", cfl);
    std::string s = b;
    return s.size() > 3 ? std::string("***") : s;
}

bool stem(const gpm_io_meta *m, std::string &out)
{
    if (!m) return false;
    const std::string p = problem_name(m->problem), me = method_name(*m);
    if (p.empty() || me.empty()) return false;
    out = "output_" + p + "_" + me + "_CFL_" + trimmed(cfl_name(m->cfl).c_str()) + "_" + std::to_string(m->lf) + "_" +
          std::to_string(m->nf);
    return true;
}

// A 1-D (or 2-D, Fortran-reversed) dataset of a native type.
bool put(hid_t f, const char *name, hid_t type, int rank, const hsize_t *dims, const void *data)
{
    const hid_t sp = H5Screate_simple(rank, dims, nullptr);
    if (sp < 0) return false;
    const hid_t ds = H5Dcreate2(f, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = ds >= 0;
    if (ok) ok = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    return ok;
}

// A string as a 1-D dataset of one-byte H5T_C_S1 elements (h5dcreate_f with
// H5T_C_S1 over dims = [len(trim(s))]).
bool put_string(hid_t f, const char *name, const std::string &s)
{
    const hid_t t = H5Tcopy(H5T_C_S1);
    if (t < 0) return false;
    H5Tset_size(t, 1);
    const hsize_t d = s.size();
    const hid_t sp = H5Screate_simple(1, &d, nullptr);
    const hid_t ds = sp >= 0 ? H5Dcreate2(f, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) : -1;
    bool ok = ds >= 0;
    if (ok && d > 0) ok = H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, s.data()) >= 0;
    if (ds >= 0) H5Dclose(ds);
    if (sp >= 0) H5Sclose(sp);
    H5Tclose(t);
    return ok;
}

bool put_int(hid_t f, const char *name, int32_t v)
{
    const hsize_t d = 1;
    return put(f, name, H5T_NATIVE_INT32, 1, &d, &v);
}

bool put_strings(hid_t f, const gpm_io_meta *m)
{
    return put_string(f, "CFL", trimmed(cfl_name(m->cfl).c_str())) && put_string(f, "problem", problem_name(m->problem)) &&
           put_string(f, "method", method_name(*m)) && put_string(f, "NN_filename", trimmed(m->nn_filename));
}

bool get(hid_t f, const char *name, hid_t type, size_t count, void *data)
{
    const hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds < 0) return false;
    const hid_t sp = H5Dget_space(ds);
    const hssize_t n = sp >= 0 ? H5Sget_simple_extent_npoints(sp) : -1;
    bool ok = n == (hssize_t)count && H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    if (sp >= 0) H5Sclose(sp);
    H5Dclose(ds);
    return ok;
}

}  // namespace

extern "C" {

int gpm_io_file_name(const gpm_io_meta *meta, int32_t niter, char *buf, int32_t bufsize)
{
    std::string s;
    if (!buf || bufsize <= 0 || !stem(meta, s)) return GPM_IO_EINVAL;
    if (niter >= 0) {
        char id[32];
        std::snprintf(id, sizeof(id), "%6d", niter + 100000);  // write(fileID, '(i6)') fileNumb + 100000
        s += "_" + trimmed(id) + ".h5";
    }
    if ((int32_t)s.size() + 1 > bufsize) return GPM_IO_EINVAL;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return GPM_IO_OK;
}

int gpm_io_write_output(const char *path, const gpm_io_meta *meta, const double *U, const int32_t *ordr,
                        const int32_t *ordr_priori, int32_t mord_if_null, int32_t niter, double t)
{
    std::string st;
    if (!path || !U || !ordr || !stem(meta, st) || meta->lf <= 0 || meta->nf <= 0) return GPM_IO_EINVAL;
    const int lf = meta->lf, nf = meta->nf, W = lf + 8;
    const size_t n = (size_t)lf * nf;
    std::vector<double> prim[4];
    for (auto &v : prim) v.resize(n);
    for (int j = 0; j < nf; ++j)
        for (int i = 0; i < lf; ++i) {
            // conservative_to_primitive (physics.f90:162-175) with pressure (:132-143)
            const double *u = U + ((size_t)(j + 4) * W + (i + 4)) * 4;
            const size_t c = (size_t)j * lf + i;
            prim[0][c] = u[0];
            prim[1][c] = u[1] / u[0];
            prim[2][c] = u[2] / u[0];
            prim[3][c] = (kGamma - 1.0) * (u[3] - (0.5 * (u[1] * u[1] + u[2] * u[2])) / u[0]);
        }
    std::vector<int32_t> prio;
    if (!ordr_priori) prio.assign(n, mord_if_null);
    const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return GPM_IO_EFILE;
    const hsize_t dims[2] = {(hsize_t)nf, (hsize_t)lf}, one = 1;
    const bool ok = put(f, "rho", H5T_NATIVE_DOUBLE, 2, dims, prim[0].data()) &&
                    put(f, "ux", H5T_NATIVE_DOUBLE, 2, dims, prim[1].data()) &&
                    put(f, "uy", H5T_NATIVE_DOUBLE, 2, dims, prim[2].data()) &&
                    put(f, "p", H5T_NATIVE_DOUBLE, 2, dims, prim[3].data()) &&
                    put(f, "ordr", H5T_NATIVE_INT32, 2, dims, ordr) &&
                    put(f, "ordr_priori", H5T_NATIVE_INT32, 2, dims, ordr_priori ? ordr_priori : prio.data()) &&
                    put_strings(f, meta) && put_int(f, "lf", lf) && put_int(f, "nf", nf) &&
                    put_int(f, "niter", niter) && put(f, "time", H5T_NATIVE_DOUBLE, 1, &one, &t);
    H5Fclose(f);
    return ok ? GPM_IO_OK : GPM_IO_EFILE;
}

int gpm_io_read_restart(const char *path, int32_t lf, int32_t nf, double *U, double *t, int32_t *niter)
{
    if (!path || !U || !t || !niter || lf <= 0 || nf <= 0) return GPM_IO_EINVAL;
    const size_t n = (size_t)lf * nf;
    std::vector<double> prim[4];
    for (auto &v : prim) v.resize(n);
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return GPM_IO_EFILE;
    double tt = 0;
    int32_t it = 0;
    const bool ok = get(f, "rho", H5T_NATIVE_DOUBLE, n, prim[0].data()) &&
                    get(f, "ux", H5T_NATIVE_DOUBLE, n, prim[1].data()) &&
                    get(f, "uy", H5T_NATIVE_DOUBLE, n, prim[2].data()) &&
                    get(f, "p", H5T_NATIVE_DOUBLE, n, prim[3].data()) && get(f, "time", H5T_NATIVE_DOUBLE, 1, &tt) &&
                    get(f, "niter", H5T_NATIVE_INT32, 1, &it);
    H5Fclose(f);
    if (!ok) return GPM_IO_EFILE;
    const int W = lf + 8;
    for (int j = 0; j < nf; ++j)
        for (int i = 0; i < lf; ++i) {
            // primitive_to_conservative (physics.f90:146-160)
            const size_t c = (size_t)j * lf + i;
            const double r = prim[0][c], vx = prim[1][c], vy = prim[2][c], p = prim[3][c];
            double *u = U + ((size_t)(j + 4) * W + (i + 4)) * 4;
            u[0] = r;
            u[1] = r * vx;
            u[2] = r * vy;
            u[3] = (p / (kGamma - 1.)) + 0.5 * r * (vx * vx + vy * vy);
        }
    *t = tt;
    *niter = it;
    return GPM_IO_OK;
}

int gpm_io_write_diagnostic(const char *path, const gpm_io_meta *meta, int32_t n, const float *time,
                            const float *pct, const float *pct_priori, const float *pct_posteriori,
                            const int32_t *steps_nn_nan, const int32_t *steps_nn_sample,
                            int32_t count_steps_nn_nan, int32_t iter_0, int32_t niter)
{
    std::string st;
    if (!path || n < 0 || !stem(meta, st) || (n > 0 && (!time || !pct || !pct_priori || !pct_posteriori)))
        return GPM_IO_EINVAL;
    std::vector<int32_t> zeros((size_t)n, 0);
    const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return GPM_IO_EFILE;
    const hsize_t d = (hsize_t)n;
    const bool ok = put(f, "time", H5T_NATIVE_FLOAT, 1, &d, time) && put(f, "pct_detected_cell", H5T_NATIVE_FLOAT, 1, &d, pct) &&
                    put(f, "pct_detected_cell_a_priori", H5T_NATIVE_FLOAT, 1, &d, pct_priori) &&
                    put(f, "pct_detected_cell_a_posteriori", H5T_NATIVE_FLOAT, 1, &d, pct_posteriori) &&
                    put(f, "steps_NN_produced_NAN", H5T_NATIVE_INT32, 1, &d, steps_nn_nan ? steps_nn_nan : zeros.data()) &&
                    put(f, "steps_NN_sample", H5T_NATIVE_INT32, 1, &d, steps_nn_sample ? steps_nn_sample : zeros.data()) &&
                    put_strings(f, meta) && put_int(f, "lf", meta->lf) && put_int(f, "nf", meta->nf) &&
                    put_int(f, "count_steps_NN_produced_NAN", count_steps_nn_nan) && put_int(f, "iter_0", iter_0) &&
                    put_int(f, "niter", niter);
    H5Fclose(f);
    return ok ? GPM_IO_OK : GPM_IO_EFILE;
}

int gpm_io_write_nn_dataset(const char *path, const gpm_io_meta *meta, int32_t size, int32_t L, const float *inputs,
                            const float *labels, int32_t n_overwrite, int32_t nr0, int32_t nr1)
{
    std::string st;
    if (!path || size < 0 || L <= 0 || !stem(meta, st) || (size > 0 && (!inputs || !labels))) return GPM_IO_EINVAL;
    const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return GPM_IO_EFILE;
    const hsize_t dl[2] = {(hsize_t)size, 2}, di[2] = {(hsize_t)size, (hsize_t)L};
    const bool ok = put(f, "labels", H5T_NATIVE_FLOAT, 2, dl, labels) && put(f, "inputs", H5T_NATIVE_FLOAT, 2, di, inputs) &&
                    put_strings(f, meta) && put_int(f, "lf", meta->lf) && put_int(f, "nf", meta->nf) &&
                    put_int(f, "n_overwrite", n_overwrite) && put_int(f, "NR0", nr0) && put_int(f, "NR1", nr1);
    H5Fclose(f);
    return ok ? GPM_IO_OK : GPM_IO_EFILE;
}

}  // extern "C"
