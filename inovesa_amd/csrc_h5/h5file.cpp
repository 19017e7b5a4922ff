// Result and checkpoint files in the layout of Inovesa's HDF5File
// (src/IO/HDF5File.cpp), over the HDF5 C library.
//
// The layout follows the reference's constructor: the groups of
// _prepareFile (:558-587), one chunked, shuffled, deflate-6 data set per
// quantity with an unlimited time dimension (_makeDatasetInfo, :514-550),
// the unit attributes and the soft links from each quantity's axis0/axis1
// to the axis data sets under /Info.  Appending extends the time dimension
// by the rows written (_appendData, :497-511).
#include <hdf5.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "vfps_hdf5.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg)
{
    g_err = msg;
    return -1;  // VFPS_ERR_INVALID
}

constexpr unsigned COMPRESSION = 6;  // HDF5File::compression

struct Dataset {
    hid_t id = -1;
    hid_t type = -1;   // file and memory type
    int rank = 0;
    std::array<hsize_t, 4> dims{};
};

}  // namespace

struct vfps_h5 {
    hid_t file = -1;
    vfps_h5_info info{};
    std::map<std::string, Dataset> ds;
};

namespace {

// _makeDatasetInfo: chunk and maximum dimensions of at least 1; a chunk
// dimension is at most its fixed maximum (HDF5 requires it)
bool make_dataset(vfps_h5* f, const std::string& name, hid_t type, int rank,
                  std::array<hsize_t, 4> dims, std::array<hsize_t, 4> chunk,
                  std::array<hsize_t, 4> maxdims)
{
    for (int i = 0; i < rank; i++) {
        chunk[i] = std::max<hsize_t>(chunk[i], 1);
        if (maxdims[i] != H5S_UNLIMITED) {
            maxdims[i] = std::max<hsize_t>(maxdims[i], 1);
            chunk[i] = std::min(chunk[i], maxdims[i]);
        }
    }
    const hid_t space = H5Screate_simple(rank, dims.data(), maxdims.data());
    const hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(dcpl, rank, chunk.data());
    H5Pset_shuffle(dcpl);
    H5Pset_deflate(dcpl, COMPRESSION);
    const hid_t id = H5Dcreate2(f->file, name.c_str(), type, space,
                                H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Pclose(dcpl);
    H5Sclose(space);
    if (id < 0) {
        return false;
    }
    Dataset d;
    d.id = id;
    d.type = type;
    d.rank = rank;
    d.dims = dims;
    f->ds[name] = d;
    return true;
}

bool attr_f64(hid_t obj, const char* name, double v)
{
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(obj, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT,
                               H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_DOUBLE, &v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

bool attr_i64(hid_t obj, const char* name, int64_t v)
{
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(obj, name, H5T_STD_I64LE, sp, H5P_DEFAULT,
                               H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_INT64, &v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

bool attr_str(hid_t obj, const char* name, const char* v)
{
    const hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, std::max<size_t>(std::string(v).size(), 1));
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, t, v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    H5Tclose(t);
    return ok;
}

bool attr_typed(hid_t obj, const char* name, hid_t ftype, hid_t mtype,
                const void* v)
{
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(obj, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, mtype, v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

bool ds_attr(vfps_h5* f, const std::string& ds, const char* name, double v)
{
    return attr_f64(f->ds.at(ds).id, name, v);
}

bool link(vfps_h5* f, const char* target, const char* name)
{
    return H5Lcreate_soft(target, f->file, name, H5P_DEFAULT, H5P_DEFAULT) >= 0;
}

bool write_all(vfps_h5* f, const std::string& name, const void* data)
{
    const Dataset& d = f->ds.at(name);
    return H5Dwrite(d.id, d.type == H5T_NATIVE_UINT32 ? H5T_NATIVE_UINT32
                                                      : H5T_NATIVE_FLOAT,
                    H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

// _appendData: `rows` more rows along the time dimension
bool append(vfps_h5* f, const std::string& name, const void* data,
            hsize_t rows = 1)
{
    Dataset& d = f->ds.at(name);
    if (rows == 0) {
        return true;
    }
    std::array<hsize_t, 4> offset{};
    offset[0] = d.dims[0];
    std::array<hsize_t, 4> ext = d.dims;
    ext[0] = rows;
    d.dims[0] += rows;
    if (H5Dset_extent(d.id, d.dims.data()) < 0) {
        return false;
    }
    const hid_t fs = H5Dget_space(d.id);
    H5Sselect_hyperslab(fs, H5S_SELECT_SET, offset.data(), nullptr,
                        ext.data(), nullptr);
    const hid_t ms = H5Screate_simple(d.rank, ext.data(), nullptr);
    const bool ok = H5Dwrite(d.id, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT,
                             data) >= 0;
    H5Sclose(ms);
    H5Sclose(fs);
    return ok;
}

const char* GROUPS[] = {
    "/BunchLength", "/BunchPopulation", "/BunchPosition", "/BunchProfile",
    "/CSR", "/CSR/Intensity", "/CSR/Spectrum", "/EnergyProfile",
    "/EnergySpread", "/EnergyAverage", "/Particles", "/PhaseSpace",
    "/RFKicks", "/Impedance", "/Impedance/data", "/Info", "/Info/Parameters",
    "/SourceMap", "/SourceMap/data", "/WakeFunction", "/WakePotential"};

bool layout(vfps_h5* f)
{
    const vfps_h5_info& in = f->info;
    const hsize_t U = H5S_UNLIMITED;
    const hid_t F = H5T_IEEE_F32LE;
    const hsize_t nb = in.nb, nx = in.nx, ny = in.ny, nbk = in.nbuckets;
    const hsize_t maxn = in.maxn, imp = in.imp_size, np = in.nparticles;
    const hsize_t c256x = std::min<hsize_t>(256, nx);
    const hsize_t c256y = std::min<hsize_t>(256, ny);
    for (const char* g : GROUPS) {
        const hid_t id = H5Gcreate2(f->file, g, H5P_DEFAULT, H5P_DEFAULT,
                                    H5P_DEFAULT);
        if (id < 0) return false;
        H5Gclose(id);
    }
    // src/IO/HDF5File.cpp:44-149
    bool ok = true;
    ok &= make_dataset(f, "/Info/AxisValues_z", F, 1, {nx},
                       {std::min<hsize_t>(2048, nx)}, {nx});
    ok &= make_dataset(f, "/Info/AxisValues_E", F, 1, {ny},
                       {std::min<hsize_t>(2048, ny)}, {ny});
    ok &= make_dataset(f, "/Info/AxisValues_f", F, 1, {maxn},
                       {std::min<hsize_t>(2048, ny)}, {maxn});
    ok &= make_dataset(f, "/Info/AxisValues_t", F, 1, {0}, {256}, {U});
    ok &= make_dataset(f, "/PhaseSpace/axis0", F, 1, {0}, {256}, {U});
    ok &= make_dataset(f, "/Info/BucketNumbers", H5T_NATIVE_UINT32, 1, {nbk},
                       {std::min<hsize_t>(2048, nbk)}, {nbk});
    ok &= make_dataset(f, "/BunchPopulation/data", F, 2, {0, nb}, {256, nb},
                       {U, nb});
    ok &= make_dataset(f, "/BunchProfile/data", F, 3, {0, nb, nx},
                       {64, 1, c256x}, {U, nb, nx});
    ok &= make_dataset(f, "/BunchProfile/padded", F, 2, {0, imp}, {2, c256x},
                       {U, imp});
    ok &= make_dataset(f, "/BunchLength/data", F, 2, {0, nb}, {256, 1},
                       {U, nb});
    ok &= make_dataset(f, "/BunchPosition/data", F, 2, {0, nb}, {256, 1},
                       {U, nb});
    ok &= make_dataset(f, "/EnergyProfile/data", F, 3, {0, nb, nx},
                       {64, 1, c256x}, {U, nb, nx});
    ok &= make_dataset(f, "/EnergySpread/data", F, 2, {0, nb}, {256, 1},
                       {U, nb});
    ok &= make_dataset(f, "/EnergyAverage/data", F, 2, {0, nb}, {256, 1},
                       {U, nb});
    ok &= make_dataset(f, "/Particles/data", F, 3, {0, np, 2},
                       {256, std::min<hsize_t>(256, np), 2}, {U, np, 2});
    ok &= make_dataset(f, "/RFKicks/data", F, 2, {0, 2}, {256, 1}, {U, 2});
    ok &= make_dataset(f, "/WakePotential/data", F, 3, {0, nb, nx},
                       {64, 1, c256x}, {U, nb, nx});
    ok &= make_dataset(f, "/WakePotential/padded", F, 2, {0, imp}, {2, c256x},
                       {U, imp});
    ok &= make_dataset(f, "/CSR/Spectrum/data", F, 3, {0, nb, maxn},
                       {64, 1, std::min<hsize_t>(256, maxn)}, {U, nb, maxn});
    ok &= make_dataset(f, "/CSR/Intensity/data", F, 2, {0, nb}, {64, 1},
                       {U, nb});
    ok &= make_dataset(f, "/PhaseSpace/data", F, 4, {0, nb, nx, ny},
                       {64, 1, c256x, c256y}, {U, nb, nx, ny});
    ok &= make_dataset(f, "/Impedance/data/real", F, 1, {imp},
                       {std::min<hsize_t>(4097, imp)}, {imp});
    ok &= make_dataset(f, "/Impedance/data/imag", F, 1, {imp},
                       {std::min<hsize_t>(4097, imp)}, {imp});
    if (!ok) return false;

    // axes (:152-186).  /Info/AxisValues_E receives the values of axis 0,
    // as the reference writes it (:168).
    const double z_meter = in.meter, z_second = in.meter / 2.99792458e8;
    ok &= ds_attr(f, "/Info/AxisValues_z", "Meter", z_meter);
    ok &= ds_attr(f, "/Info/AxisValues_z", "Second", z_second);
    ok &= write_all(f, "/Info/AxisValues_z", in.axis_z);
    ok &= ds_attr(f, "/Info/AxisValues_E", "ElectronVolt", in.electronvolt);
    ok &= write_all(f, "/Info/AxisValues_E", in.axis_z);
    if (maxn > 0 && in.axis_f) {
        ok &= ds_attr(f, "/Info/AxisValues_f", "Hertz", in.hertz);
        ok &= write_all(f, "/Info/AxisValues_f", in.axis_f);
    }
    ok &= ds_attr(f, "/Info/AxisValues_t", "Second", in.t_sync);
    ok &= ds_attr(f, "/Info/AxisValues_t", "Turn", in.t_sync * in.f_rev);
    ok &= ds_attr(f, "/BunchPopulation/data", "Ampere", in.current);
    ok &= ds_attr(f, "/BunchPopulation/data", "Coulomb", in.charge);
    if (nbk > 0) {
        ok &= write_all(f, "/Info/BucketNumbers", in.buckets);
    }
    // quantities, their units and axes (:196-305)
    ok &= link(f, "/Info/AxisValues_t", "/BunchProfile/axis0");
    ok &= link(f, "/Info/AxisValues_z", "/BunchProfile/axis1");
    ok &= ds_attr(f, "/BunchProfile/data", "AmperePerNBL", in.current);
    ok &= ds_attr(f, "/BunchProfile/data", "CoulombPerNBL", in.charge);
    ok &= link(f, "/Info/AxisValues_t", "/BunchLength/axis0");
    ok &= link(f, "/Info/AxisValues_t", "/BunchPopulation/axis0");
    ok &= ds_attr(f, "/BunchLength/data", "Meter", z_meter);
    ok &= ds_attr(f, "/BunchLength/data", "Second", z_second);
    ok &= link(f, "/Info/AxisValues_t", "/BunchPosition/axis0");
    ok &= ds_attr(f, "/BunchPosition/data", "Meter", z_meter);
    ok &= ds_attr(f, "/BunchPosition/data", "Second", z_second);
    ok &= link(f, "/Info/AxisValues_t", "/EnergyProfile/axis0");
    ok &= link(f, "/Info/AxisValues_E", "/EnergyProfile/axis1");
    ok &= ds_attr(f, "/EnergyProfile/data", "AmperePerNES", in.current);
    ok &= ds_attr(f, "/EnergyProfile/data", "CoulombPerNES", in.charge);
    ok &= ds_attr(f, "/EnergySpread/data", "ElectronVolt", in.electronvolt);
    ok &= link(f, "/Info/AxisValues_t", "/EnergyAverage/axis0");
    ok &= link(f, "/Info/AxisValues_t", "/EnergySpread/axis0");
    ok &= ds_attr(f, "/EnergyAverage/data", "ElectronVolt", in.electronvolt);
    ok &= link(f, "/Info/AxisValues_t", "/RFKicks/axis0");
    ok &= link(f, "/Info/AxisValues_t", "/Particles/axis0");
    if (nbk > 0) {
        ok &= link(f, "/Info/AxisValues_t", "/WakePotential/axis0");
        ok &= link(f, "/Info/AxisValues_z", "/WakePotential/axis1");
        ok &= ds_attr(f, "/WakePotential/data", "Volt", in.volts);
        ok &= link(f, "/Info/AxisValues_t", "/CSR/Spectrum/axis0");
        ok &= link(f, "/Info/AxisValues_f", "/CSR/Spectrum/axis1");
        ok &= ds_attr(f, "/CSR/Spectrum/data", "WattPerHertz",
                      in.factor4WattPerHertz);
        ok &= link(f, "/Info/AxisValues_t", "/CSR/Intensity/axis0");
        ok &= ds_attr(f, "/CSR/Intensity/data", "Watt", in.factor4Watts);
    }
    // the phase space's own time axis (:307-324)
    ok &= ds_attr(f, "/PhaseSpace/axis0", "Second", in.t_sync);
    ok &= ds_attr(f, "/PhaseSpace/axis0", "Turn", in.t_sync * in.f_rev);
    ok &= link(f, "/Info/AxisValues_z", "/PhaseSpace/axis1");
    ok &= link(f, "/Info/AxisValues_E", "/PhaseSpace/axis2");
    ok &= ds_attr(f, "/PhaseSpace/data", "AmperePerNBLPerNES", in.current);
    ok &= ds_attr(f, "/PhaseSpace/data", "CoulombPerNBLPerNES", in.charge);
    // impedance (:327-346)
    if (imp > 0 && in.imp_real && in.imp_imag) {
        ok &= link(f, "/Info/AxisValues_f", "/Impedance/axis0");
        ok &= write_all(f, "/Impedance/data/real", in.imp_real);
        ok &= write_all(f, "/Impedance/data/imag", in.imp_imag);
        const hid_t g = H5Gopen2(f->file, "/Impedance/data", H5P_DEFAULT);
        ok &= g >= 0 && attr_f64(g, "Ohm", in.factor4Ohms);
        if (g >= 0) H5Gclose(g);
    }
    // version (:348-377)
    {
        const hsize_t three = 3;
        const hid_t sp = H5Screate_simple(1, &three, &three);
        const hid_t d = H5Dcreate2(f->file, "/Info/Inovesa_v", H5T_STD_I32LE,
                                   sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        ok &= d >= 0 && H5Dwrite(d, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL,
                                 H5P_DEFAULT, in.version) >= 0;
        if (d >= 0) H5Dclose(d);
        H5Sclose(sp);
    }
    {
        const std::string b = in.build ? in.build : "";
        const hsize_t len = std::max<size_t>(b.size(), 1);
        const hid_t sp = H5Screate_simple(1, &len, nullptr);
        const hid_t d = H5Dcreate2(f->file, "/Info/Inovesa_build", H5T_C_S1,
                                   sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        std::string buf = b.empty() ? std::string(1, '\0') : b;
        ok &= d >= 0 && H5Dwrite(d, H5T_C_S1, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                 buf.data()) >= 0;
        if (d >= 0) H5Dclose(d);
        H5Sclose(sp);
    }
    return ok;
}

void close_all(vfps_h5* f)
{
    for (auto& kv : f->ds) {
        if (kv.second.id >= 0) H5Dclose(kv.second.id);
    }
    f->ds.clear();
    if (f->file >= 0) {
        H5Fclose(f->file);
        f->file = -1;
    }
}

}  // namespace

extern "C" {

const char* vfps_h5_last_error(void) { return g_err.c_str(); }

int vfps_h5_create(const char* path, const vfps_h5_info* info, vfps_h5** out)
{
    if (!path || !info || !out || info->nb == 0 || info->nx == 0 ||
        info->ny == 0 || !info->axis_z) {
        return fail("vfps_h5_create: bad argument");
    }
    if (info->nbuckets && !info->buckets) {
        return fail("vfps_h5_create: bucket numbers missing");
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    auto* f = new vfps_h5();
    f->info = *info;
    f->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f->file < 0) {
        delete f;
        return fail(std::string("vfps_h5_create: cannot create ") + path);
    }
    if (!layout(f)) {
        close_all(f);
        delete f;
        return fail("vfps_h5_create: HDF5 error while laying out the file");
    }
    *out = f;
    return 0;
}

int vfps_h5_close(vfps_h5* f)
{
    if (!f) return 0;
    close_all(f);
    delete f;
    return 0;
}

int vfps_h5_append_ps(vfps_h5* f, float t, int at, const float* mesh,
                      const float* xproj, const float* yproj,
                      const float* length, const float* position,
                      const float* spread, const float* average,
                      const float* population)
{
    if (!f || at < VFPS_H5_ALL || at > VFPS_H5_PHASESPACE) {
        return fail("append: bad argument");
    }
    bool ok = true;
    if (at == VFPS_H5_ALL || at == VFPS_H5_PHASESPACE) {
        if (!mesh) return fail("append: phase space missing");
        ok &= append(f, "/PhaseSpace/axis0", &t);
        ok &= append(f, "/PhaseSpace/data", mesh);
    }
    if (at != VFPS_H5_PHASESPACE) {
        if (!xproj || !yproj || !length || !position || !spread || !average ||
            !population) {
            return fail("append: projections or moments missing");
        }
        ok &= append(f, "/Info/AxisValues_t", &t);
        ok &= append(f, "/BunchProfile/data", xproj);
        ok &= append(f, "/BunchLength/data", length);
        ok &= append(f, "/BunchPosition/data", position);
        ok &= append(f, "/EnergyProfile/data", yproj);
        ok &= append(f, "/EnergySpread/data", spread);
        ok &= append(f, "/EnergyAverage/data", average);
        ok &= append(f, "/BunchPopulation/data", population);
    }
    return ok ? 0 : fail("append: HDF5 write failed");
}

int vfps_h5_append_csr(vfps_h5* f, const float* power, const float* spectrum)
{
    if (!f || !power) return fail("append(ElectricField): bad argument");
    bool ok = true;
    if (spectrum) ok &= append(f, "/CSR/Spectrum/data", spectrum);
    ok &= append(f, "/CSR/Intensity/data", power);
    return ok ? 0 : fail("append(ElectricField): HDF5 write failed");
}

int vfps_h5_append_padded(vfps_h5* f, const float* profile, const float* wake)
{
    if (!f || !profile || !wake) return fail("appendPadded: bad argument");
    bool ok = append(f, "/BunchProfile/padded", profile);
    ok &= append(f, "/WakePotential/padded", wake);
    return ok ? 0 : fail("appendPadded: HDF5 write failed");
}

int vfps_h5_append_rfkicks(vfps_h5* f, const float* kicks, uint32_t count)
{
    if (!f || (count && !kicks)) return fail("appendRFKicks: bad argument");
    return append(f, "/RFKicks/data", kicks, count)
               ? 0
               : fail("appendRFKicks: HDF5 write failed");
}

int vfps_h5_append_tracks(vfps_h5* f, const float* positions)
{
    if (!f || (f->info.nparticles && !positions)) {
        return fail("appendTracks: bad argument");
    }
    return append(f, "/Particles/data", positions)
               ? 0
               : fail("appendTracks: HDF5 write failed");
}

int vfps_h5_append_wake(vfps_h5* f, const float* wake)
{
    if (!f || !wake) return fail("append(WakeKickMap): bad argument");
    return append(f, "/WakePotential/data", wake)
               ? 0
               : fail("append(WakeKickMap): HDF5 write failed");
}

static int add_parameter(vfps_h5* f, const char* group, const char* name,
                         int kind, double d, int64_t i, const char* s)
{
    if (!f || !group || !name) return fail("addParameterToGroup: bad argument");
    const hid_t g = H5Gopen2(f->file, group, H5P_DEFAULT);
    if (g < 0) return fail(std::string("addParameterToGroup: no group ") + group);
    bool ok = kind == 0 ? attr_f64(g, name, d)
                        : (kind == 1 ? attr_i64(g, name, i)
                                     : attr_str(g, name, s ? s : ""));
    H5Gclose(g);
    return ok ? 0 : fail(std::string("addParameterToGroup: ") + name);
}

int vfps_h5_add_parameter_f64(vfps_h5* f, const char* group, const char* name,
                              double value)
{
    return add_parameter(f, group, name, 0, value, 0, nullptr);
}

int vfps_h5_add_parameter_i64(vfps_h5* f, const char* group, const char* name,
                              int64_t value)
{
    return add_parameter(f, group, name, 1, 0.0, value, nullptr);
}

int vfps_h5_add_parameter_str(vfps_h5* f, const char* group, const char* name,
                              const char* value)
{
    return add_parameter(f, group, name, 2, 0.0, 0, value);
}

int vfps_h5_add_parameter_typed(vfps_h5* f, const char* group,
                                const char* name, int type, const void* value)
{
    if (!f || !group || !name || !value) {
        return fail("addParameterToGroup: bad argument");
    }
    hid_t ft, mt;
    switch (type) {
    case VFPS_H5_F32: ft = H5T_IEEE_F32LE; mt = H5T_NATIVE_FLOAT; break;
    case VFPS_H5_F64: ft = H5T_IEEE_F64LE; mt = H5T_NATIVE_DOUBLE; break;
    case VFPS_H5_U32: ft = H5T_STD_U32LE; mt = H5T_NATIVE_UINT32; break;
    case VFPS_H5_I32: ft = H5T_STD_I32LE; mt = H5T_NATIVE_INT32; break;
    default: return fail("addParameterToGroup: type");
    }
    const hid_t g = H5Gopen2(f->file, group, H5P_DEFAULT);
    if (g < 0) return fail(std::string("addParameterToGroup: no group ") + group);
    const bool ok = attr_typed(g, name, ft, mt, value);
    H5Gclose(g);
    return ok ? 0 : fail(std::string("addParameterToGroup: ") + name);
}

int vfps_h5_rows(vfps_h5* f, const char* dataset, uint64_t* rows)
{
    if (!f || !dataset || !rows) return fail("vfps_h5_rows: bad argument");
    auto it = f->ds.find(dataset);
    if (it == f->ds.end()) return fail(std::string("no data set ") + dataset);
    *rows = it->second.dims[0];
    return 0;
}

int vfps_h5_read_phasespace(const char* path, int64_t step, uint32_t* nb,
                            uint32_t* n, uint64_t* nsteps, float* out)
{
    if (!path) return fail("readPhaseSpace: NULL path");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return fail(std::string("readPhaseSpace: cannot open ") + path);
    const hid_t d = H5Dopen2(file, "/PhaseSpace/data", H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(file);
        return fail("readPhaseSpace: no /PhaseSpace/data");
    }
    const hid_t fs = H5Dget_space(d);
    const int rank = H5Sget_simple_extent_ndims(fs);
    std::array<hsize_t, 4> dims{};
    int rc = 0;
    if (rank != 3 && rank != 4) {
        rc = fail("readPhaseSpace: unexpected rank");
    } else {
        H5Sget_simple_extent_dims(fs, dims.data(), nullptr);
        const uint32_t nbunch = rank == 4 ? (uint32_t)dims[1] : 1u;
        const hsize_t size = dims[rank == 4 ? 2 : 1];
        const hsize_t size2 = dims[rank == 4 ? 3 : 2];
        if (nb) *nb = nbunch;
        if (n) *n = (uint32_t)size;
        if (nsteps) *nsteps = dims[0];
        if (size != size2) {
            rc = fail("readPhaseSpace: the phase space is not square");
        } else if (dims[0] == 0) {
            rc = fail("readPhaseSpace: no phase space saved");
        } else if (out) {
            // use_step = (ps_dims[0] + use_step) % ps_dims[0] (:452)
            const int64_t T = (int64_t)dims[0];
            const int64_t use = ((T + step) % T + T) % T;
            std::array<hsize_t, 4> offset{};
            std::array<hsize_t, 4> ext = dims;
            offset[0] = (hsize_t)use;
            ext[0] = 1;
            H5Sselect_hyperslab(fs, H5S_SELECT_SET, offset.data(), nullptr,
                                ext.data(), nullptr);
            const hid_t ms = H5Screate_simple(rank, ext.data(), nullptr);
            if (H5Dread(d, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, out) < 0) {
                rc = fail("readPhaseSpace: read failed");
            }
            H5Sclose(ms);
        }
    }
    H5Sclose(fs);
    H5Dclose(d);
    H5Fclose(file);
    return rc;
}

}  // extern "C"
