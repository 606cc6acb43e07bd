// gam_io.cpp -- libgamera_io.so: the reference's HDF5 grid and restart files
// for one GAMERA block, on the host (HDF5 C API).  Declared in
// include/gamera_io.h.
//
// The reference writes through the HDF5 Fortran API, which reverses the
// dimension order: a Fortran array Q(n1, n2, ..., nr) becomes a dataset whose
// C dimensions are (nr, ..., n2, n1) with the bytes in Fortran order.  This
// file keeps that: a dataset is written from / read into a Fortran-ordered
// buffer as is, with its dimensions listed last to first.
//   datasets: H5LTmake_dataset_double, "Units" = "CODE" and "Description" =
//             "DESCRIPTION" string attributes (base/ioH5.F90:811-871
//             WriteHDFVar with ClearIO's defaults, :1135-1145)
//   scalars:  root attributes, H5T_NATIVE_INT / H5T_NATIVE_DOUBLE, scalar
//             dataspace (base/ioH5.F90:1185-1237 writeInt2HDF, writeReal2HDF)
//   stamp:    GITHASH, GITBRANCH, COMPILER, COMPILEROPTS, DATETIME string
//             attributes on a new file's root (base/ioH5.F90:39-69 StampIO)
#include "gamera_io.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

const double kTiny = 1.0e-12;   // base/kdefs.F90:20 TINY

// HDF5's own error stack printing is off while a call of this file runs.
struct QuietH5 {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    QuietH5() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

struct File {
    hid_t id = -1;
    ~File() { if (id >= 0) H5Fclose(id); }
};

struct Block {
    int nG, Nip, Njp, Nkp;
    hsize_t Ni() const { return (hsize_t)(Nip + 2 * nG); }
    hsize_t Nj() const { return (hsize_t)(Njp + 2 * nG); }
    hsize_t Nk() const { return (hsize_t)(Nkp + 2 * nG); }
};

bool bad_block(const Block &b) { return b.nG < 0 || b.Nip < 1 || b.Njp < 1 || b.Nkp < 1; }

// Fortran shape (n1..nr) -> C dimensions (nr..n1)
std::vector<hsize_t> cdims(std::initializer_list<hsize_t> fshape) {
    std::vector<hsize_t> d(fshape.begin(), fshape.end());
    return std::vector<hsize_t>(d.rbegin(), d.rend());
}

int set_string_attr(hid_t loc, const char *obj, const char *name, const char *value) {
    // Fortran strings: space padded, length = the trimmed text (at least 1)
    size_t n = strlen(value);
    while (n > 0 && value[n - 1] == ' ') n--;
    const size_t sz = n > 0 ? n : 1;
    std::string v(value, n);
    v.resize(sz, ' ');
    hid_t t = H5Tcopy(H5T_FORTRAN_S1);
    if (t < 0 || H5Tset_size(t, sz) < 0) return fail("string type for attribute %s", name);
    hid_t o = H5Oopen(loc, obj, H5P_DEFAULT);
    hid_t s = H5Screate(H5S_SCALAR);
    hid_t a = (o >= 0 && s >= 0) ? H5Acreate2(o, name, t, s, H5P_DEFAULT, H5P_DEFAULT) : -1;
    herr_t e = a >= 0 ? H5Awrite(a, t, v.data()) : -1;
    if (a >= 0) H5Aclose(a);
    if (s >= 0) H5Sclose(s);
    if (o >= 0) H5Oclose(o);
    H5Tclose(t);
    return e < 0 ? fail("cannot write attribute %s/%s", obj, name) : 0;
}

template <typename T>
int set_scalar_attr(hid_t file, const char *name, hid_t type, T value) {
    hid_t s = H5Screate(H5S_SCALAR);
    hid_t a = s >= 0 ? H5Acreate2(file, name, type, s, H5P_DEFAULT, H5P_DEFAULT) : -1;
    herr_t e = a >= 0 ? H5Awrite(a, type, &value) : -1;
    if (a >= 0) H5Aclose(a);
    if (s >= 0) H5Sclose(s);
    return e < 0 ? fail("cannot write attribute %s", name) : 0;
}

bool has_attr(hid_t file, const char *name) { return H5Aexists(file, name) > 0; }

template <typename T>
int get_scalar_attr(hid_t file, const char *name, hid_t type, T *out) {
    hid_t a = H5Aopen(file, name, H5P_DEFAULT);
    if (a < 0) return fail("attribute %s missing", name);
    herr_t e = H5Aread(a, type, out);
    H5Aclose(a);
    return e < 0 ? fail("cannot read attribute %s", name) : 0;
}

int write_dataset(hid_t file, const char *name, const std::vector<hsize_t> &dims, const double *data) {
    if (H5LTmake_dataset_double(file, name, (int)dims.size(), dims.data(), data) < 0)
        return fail("cannot write dataset %s", name);
    if (set_string_attr(file, name, "Units", "CODE") || set_string_attr(file, name, "Description", "DESCRIPTION"))
        return -1;
    return 0;
}

bool has_dataset(hid_t file, const char *name) {
    return H5Lexists(file, name, H5P_DEFAULT) > 0 && H5LTfind_dataset(file, name) > 0;
}

// C dimensions of dataset name (rank <= 8)
int dataset_dims(hid_t file, const char *name, std::vector<hsize_t> &dims) {
    int rank = 0;
    if (!has_dataset(file, name)) return fail("dataset %s missing", name);
    if (H5LTget_dataset_ndims(file, name, &rank) < 0 || rank < 1 || rank > 8)
        return fail("dataset %s: bad rank", name);
    dims.assign(rank, 0);
    H5T_class_t cls;
    size_t tsz;
    if (H5LTget_dataset_info(file, name, dims.data(), &cls, &tsz) < 0) return fail("dataset %s: no info", name);
    return 0;
}

int read_dataset(hid_t file, const char *name, double *out) {
    return H5LTread_dataset_double(file, name, out) < 0 ? fail("cannot read dataset %s", name) : 0;
}

int stamp(hid_t file) {
    // the reference stamps the git hash and branch of its build; this library
    // names itself there
    char when[64];
    time_t now = time(nullptr);
    struct tm tmv;
    localtime_r(&now, &tmv);
    strftime(when, sizeof when, "%Y-%m-%d %H:%M:%S", &tmv);
    if (set_string_attr(file, ".", "GITHASH", "kaiju_amd") || set_string_attr(file, ".", "GITBRANCH", "kaiju_amd") ||
        set_string_attr(file, ".", "COMPILER", "libgamera_io (HDF5 C API)") ||
        set_string_attr(file, ".", "COMPILEROPTS", "") || set_string_attr(file, ".", "DATETIME", when))
        return -1;
    return 0;
}

int create(const char *path, File &f) {
    // CheckAndKill (output.F90:157): an existing file of that name is replaced
    f.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f.id < 0) return fail("cannot create %s", path);
    return stamp(f.id);
}

int open_ro(const char *path, File &f) {
    f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    return f.id < 0 ? fail("cannot open %s", path) : 0;
}

int write_corners(hid_t file, const Block &b, const double *xyz) {
    const hsize_t n = (b.Ni() + 1) * (b.Nj() + 1) * (b.Nk() + 1);
    const auto d = cdims({b.Ni() + 1, b.Nj() + 1, b.Nk() + 1});
    if (write_dataset(file, "X", d, xyz) || write_dataset(file, "Y", d, xyz + n) ||
        write_dataset(file, "Z", d, xyz + 2 * n))
        return -1;
    return 0;
}

// active dims of the X dataset (GridSizeH5 + gioH5.F90:44-52)
int corner_active(hid_t file, int nG, int *act) {
    std::vector<hsize_t> d;
    if (dataset_dims(file, "X", d)) return -1;
    if (d.size() > 3) return fail("grid X has rank %d > 3", (int)d.size());
    long f[3] = {0, 0, 0};   // Fortran dims, 0 past the rank
    for (size_t q = 0; q < d.size(); q++) f[q] = (long)d[d.size() - 1 - q];
    for (int q = 0; q < 3; q++) act[q] = (int)(f[q] - 2L * nG - 1);
    return 0;
}

// Gas-like array: full (Fortran n1 = Ni) or active-only (n1 = Nip, older
// files) into the ghost-inclusive buffer out of ncomp components
int read_cells(hid_t file, const char *name, const Block &b, int ncomp, bool ghosts, bool face, double *out) {
    const hsize_t e = face ? 1 : 0;
    const hsize_t ni = (ghosts ? b.Ni() : (hsize_t)b.Nip) + e, nj = (ghosts ? b.Nj() : (hsize_t)b.Njp) + e,
                  nk = (ghosts ? b.Nk() : (hsize_t)b.Nkp) + e;
    std::vector<hsize_t> d;
    if (dataset_dims(file, name, d)) return -1;
    hsize_t total = 1;
    for (hsize_t x : d) total *= x;
    if (total != ni * nj * nk * (hsize_t)ncomp || d.back() != ni)
        return fail("dataset %s has the shape of another block", name);
    if (ghosts) return read_dataset(file, name, out);
    std::vector<double> tmp(total);
    if (read_dataset(file, name, tmp.data())) return -1;
    const hsize_t Ni = b.Ni() + e, Nj = b.Nj() + e, Nk = b.Nk() + e, g = (hsize_t)b.nG;
    for (int v = 0; v < ncomp; v++)
        for (hsize_t k = 0; k < nk; k++)
            for (hsize_t j = 0; j < nj; j++)
                for (hsize_t i = 0; i < ni; i++)
                    out[(((hsize_t)v * Nk + k + g) * Nj + j + g) * Ni + i + g] = tmp[(((hsize_t)v * nk + k) * nj + j) * ni + i];
    return 0;
}

}  // namespace

extern "C" const char *gam_io_last_error(void) { return g_err.c_str(); }

extern "C" int gam_io_grid_dims(const char *path, int nG, int *dims) {
    if (!path || !dims || nG < 0) return fail("gam_io_grid_dims: bad argument");
    QuietH5 q;
    File f;
    if (open_ro(path, f)) return -1;
    return corner_active(f.id, nG, dims);
}

extern "C" int gam_io_read_grid(const char *path, int nG, int Nip, int Njp, int Nkp, double *xyz) {
    const Block b{nG, Nip, Njp, Nkp};
    if (!path || !xyz || bad_block(b)) return fail("gam_io_read_grid: bad argument");
    QuietH5 q;
    File f;
    if (open_ro(path, f)) return -1;
    int act[3];
    if (corner_active(f.id, nG, act)) return -1;
    if (act[0] != Nip || act[1] != Njp || act[2] != Nkp)
        return fail("%s: grid of %d x %d x %d active cells, not %d x %d x %d", path, act[0], act[1], act[2], Nip, Njp, Nkp);
    const hsize_t n = (b.Ni() + 1) * (b.Nj() + 1) * (b.Nk() + 1);
    if (read_dataset(f.id, "X", xyz) || read_dataset(f.id, "Y", xyz + n) || read_dataset(f.id, "Z", xyz + 2 * n))
        return -1;
    return 0;
}

extern "C" int gam_io_write_grid(const char *path, int nG, int Nip, int Njp, int Nkp, const double *xyz) {
    const Block b{nG, Nip, Njp, Nkp};
    if (!path || !xyz || bad_block(b)) return fail("gam_io_write_grid: bad argument");
    QuietH5 q;
    File f;
    if (create(path, f)) return -1;
    return write_corners(f.id, b, xyz);
}

extern "C" int gam_io_write_restart(const char *path, int nG, int Nip, int Njp, int Nkp, int doMHD,
                                    const gam_io_clock *c, const double *xyz,
                                    const double *Gas, const double *magFlux, const double *Bxyz,
                                    const double *oGas, const double *omagFlux, const double *oBxyz) {
    const Block b{nG, Nip, Njp, Nkp};
    if (!path || !c || !xyz || !Gas || !oGas || bad_block(b)) return fail("gam_io_write_restart: bad argument");
    if (doMHD && (!magFlux || !Bxyz || !omagFlux || !oBxyz)) return fail("gam_io_write_restart: MHD arrays missing");
    QuietH5 q;
    File f;
    if (create(path, f)) return -1;
    const hid_t r = f.id;
    // gioH5.F90:659-667, in the reference's order
    if (set_scalar_attr(r, "nOut", H5T_NATIVE_INT, c->nOut) || set_scalar_attr(r, "nRes", H5T_NATIVE_INT, c->nRes) ||
        set_scalar_attr(r, "ts", H5T_NATIVE_INT, c->ts) || set_scalar_attr(r, "dt", H5T_NATIVE_DOUBLE, c->dt) ||
        set_scalar_attr(r, "dt0", H5T_NATIVE_DOUBLE, c->dt0 < kTiny * 10.0 ? 0.0 : c->dt0))
        return -1;
    if (write_corners(r, b, xyz)) return -1;
    const auto dG = cdims({b.Ni(), b.Nj(), b.Nk(), 5, 1});
    const auto dM = cdims({b.Ni() + 1, b.Nj() + 1, b.Nk() + 1, 3});
    const auto dB = cdims({b.Ni(), b.Nj(), b.Nk(), 3});
    struct St { const char *id; double time; const double *G, *M, *B; };
    const St states[2] = {{"", c->t, Gas, magFlux, Bxyz}, {"o", c->ot, oGas, omagFlux, oBxyz}};
    for (const St &s : states) {   // AddState2IO, gioH5.F90:690-703
        const std::string id(s.id);
        if (set_scalar_attr(r, (id + "t").c_str(), H5T_NATIVE_DOUBLE, s.time) ||
            write_dataset(r, (id + "Gas").c_str(), dG, s.G))
            return -1;
        if (doMHD && (write_dataset(r, (id + "magFlux").c_str(), dM, s.M) || write_dataset(r, (id + "Bxyz").c_str(), dB, s.B)))
            return -1;
    }
    if (H5Fflush(r, H5F_SCOPE_GLOBAL) < 0) return fail("cannot flush %s", path);
    return 0;
}

extern "C" int gam_io_read_restart(const char *path, int nG, int Nip, int Njp, int Nkp, int doMHD, gam_io_clock *c,
                                   double *Gas, double *magFlux, double *oGas, double *omagFlux) {
    const Block b{nG, Nip, Njp, Nkp};
    if (!path || !c || !Gas || !oGas || bad_block(b)) return fail("gam_io_read_restart: bad argument");
    if (doMHD && (!magFlux || !omagFlux)) return fail("gam_io_read_restart: MHD arrays missing");
    QuietH5 q;
    File f;
    if (open_ro(path, f)) return -1;
    const hid_t r = f.id;
    if (!has_dataset(r, "oGas")) return fail("%s: restart file too old, does not include oState", path);
    std::vector<hsize_t> d;
    if (dataset_dims(r, "Gas", d)) return -1;
    if (d.size() != 5) return fail("%s: Gas has rank %d, not 5", path, (int)d.size());
    // C dims (nSpc+1, NVAR, Nk, Nj, Ni)
    if (d[0] != 1) return fail("%s: restart has %d species, this block none", path, (int)d[0] - 1);
    if (d[1] != 5) return fail("%s: Gas has %d variables, not 5", path, (int)d[1]);
    bool ghosts;
    if (d[4] == b.Ni()) ghosts = true;
    else if (d[4] == (hsize_t)b.Nip) ghosts = false;   // gioH5.F90:790-798
    else return fail("%s: restart data written from a grid of another size", path);
    gam_io_clock in{};
    if (get_scalar_attr(r, "nOut", H5T_NATIVE_INT, &in.nOut) || get_scalar_attr(r, "nRes", H5T_NATIVE_INT, &in.nRes) ||
        get_scalar_attr(r, "ts", H5T_NATIVE_INT, &in.ts) || get_scalar_attr(r, "t", H5T_NATIVE_DOUBLE, &in.t) ||
        get_scalar_attr(r, "dt", H5T_NATIVE_DOUBLE, &in.dt) || get_scalar_attr(r, "ot", H5T_NATIVE_DOUBLE, &in.ot))
        return -1;
    in.dt0 = 0.0;
    if (has_attr(r, "dt0")) {   // gioH5.F90:838-845
        if (get_scalar_attr(r, "dt0", H5T_NATIVE_DOUBLE, &in.dt0)) return -1;
        if (in.dt0 < kTiny * 10) in.dt0 = 0.0;
    }
    in.nRes += 1;   // gioH5.F90:812
    if (read_cells(r, "Gas", b, 5, ghosts, false, Gas) || read_cells(r, "oGas", b, 5, ghosts, false, oGas)) return -1;
    if (doMHD && (read_cells(r, "magFlux", b, 3, ghosts, true, magFlux) ||
                  read_cells(r, "omagFlux", b, 3, ghosts, true, omagFlux)))
        return -1;
    *c = in;
    return 0;
}

// ------------------------------------------------------------ solar-wind file
namespace {

const char *const kWindSeries[9] = {"T", "D", "Vx", "Vy", "Vz", "Temp", "Bx", "By", "Bz"};
const char *const kWindCoef[3] = {"ByC", "BzC", "Bx0"};
const char *const kWindBS[3] = {"xBS", "yBS", "zBS"};

// elements of a dataset (any rank)
int dataset_len(hid_t file, const char *name, hsize_t *n) {
    std::vector<hsize_t> d;
    if (dataset_dims(file, name, d)) return -1;
    *n = 1;
    for (hsize_t x : d) *n *= x;
    return 0;
}

// ioExist (base/ioH5.F90:348-399): a dataset or a root attribute of that name
bool io_exist(hid_t file, const char *name) { return has_dataset(file, name) || has_attr(file, name); }

// a scalar by ReadVars' rule (base/ioH5.F90:449-484): the root attribute,
// then the dataset (its first element) when there is one
int read_scalar(hid_t file, const char *name, double *out, int *found) {
    *found = 0;
    if (has_attr(file, name)) {
        if (get_scalar_attr(file, name, H5T_NATIVE_DOUBLE, out)) return -1;
        *found = 1;
    }
    if (has_dataset(file, name)) {
        hsize_t n = 0;
        if (dataset_len(file, name, &n)) return -1;
        std::vector<double> tmp(n ? n : 1);
        if (n == 0 || read_dataset(file, name, tmp.data())) return fail("cannot read %s", name);
        *out = tmp[0];
        *found = 1;
    }
    return 0;
}

int read_series(hid_t file, const char *name, int nt, double *out) {
    hsize_t n = 0;
    if (dataset_len(file, name, &n)) return -1;
    if (n != (hsize_t)nt) return fail("wind dataset %s has %d samples, T has %d", name, (int)n, nt);
    return read_dataset(file, name, out);
}

}  // namespace

extern "C" int gam_io_wind_len(const char *path, int *nt) {
    if (!path || !nt) return fail("gam_io_wind_len: bad argument");
    QuietH5 q;
    File f;
    if (open_ro(path, f)) return -1;
    hsize_t n = 0;
    if (dataset_len(f.id, "T", &n)) return -1;
    *nt = (int)n;
    return 0;
}

extern "C" int gam_io_read_wind(const char *path, int nt, double *T, double *D, double *V, double *Temp, double *B,
                                double *xyzBS, gam_io_wind_info *info) {
    if (!path || nt < 1 || !T || !D || !V || !Temp || !B || !xyzBS || !info) return fail("gam_io_read_wind: bad argument");
    QuietH5 q;
    File f;
    if (open_ro(path, f)) return -1;
    const hid_t r = f.id;
    if (!io_exist(r, "Temp"))   // wind.F90:641-646
        return fail("%s: no Temp variable (solar wind temperature rather than thermal pressure is required)", path);
    double *dst[9] = {T, D, V, V + nt, V + 2 * (size_t)nt, Temp, B, B + nt, B + 2 * (size_t)nt};
    for (int v = 0; v < 9; v++)
        if (read_series(r, kWindSeries[v], nt, dst[v])) return -1;
    gam_io_wind_info in{};
    double *coef[3] = {&in.ByC, &in.BzC, &in.Bx0};
    for (int c = 0; c < 3; c++) {   // wind.F90:704-716, absent ones 0
        *coef[c] = 0.0;
        if (read_scalar(r, kWindCoef[c], coef[c], &in.hasCoef[c])) return -1;
        if (!in.hasCoef[c]) *coef[c] = 0.0;
    }
    in.hasBS = io_exist(r, "xBS") ? 1 : 0;   // wind.F90:723-733, else the origin
    for (int x = 0; x < 3; x++) {
        double *o = xyzBS + (size_t)nt * x;
        if (in.hasBS) {
            if (read_series(r, kWindBS[x], nt, o)) return -1;
        } else {
            for (int n = 0; n < nt; n++) o[n] = 0.0;
        }
    }
    *info = in;
    return 0;
}

extern "C" int gam_io_write_wind(const char *path, int nt, const double *T, const double *D, const double *V,
                                 const double *Temp, const double *B, const double *xyzBS, const double *coef) {
    if (!path || nt < 1 || !T || !D || !V || !Temp || !B) return fail("gam_io_write_wind: bad argument");
    QuietH5 q;
    File f;
    if (create(path, f)) return -1;
    const hid_t r = f.id;
    const std::vector<hsize_t> d{(hsize_t)nt};
    const double *src[9] = {T, D, V, V + nt, V + 2 * (size_t)nt, Temp, B, B + nt, B + 2 * (size_t)nt};
    for (int v = 0; v < 9; v++)
        if (write_dataset(r, kWindSeries[v], d, src[v])) return -1;
    if (coef)
        for (int c = 0; c < 3; c++)
            if (set_scalar_attr(r, kWindCoef[c], H5T_NATIVE_DOUBLE, coef[c])) return -1;
    if (xyzBS)
        for (int x = 0; x < 3; x++)
            if (write_dataset(r, kWindBS[x], d, xyzBS + (size_t)nt * x)) return -1;
    if (H5Fflush(r, H5F_SCOPE_GLOBAL) < 0) return fail("cannot flush %s", path);
    return 0;
}
