// h5_io.cpp -- HDF5 output and restart input (see h5_io.h).
#include "h5_io.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

#ifdef TLBM_WITH_HDF5
#include "H5Cpp.h"
#endif

namespace tlbm_host {

#ifdef TLBM_WITH_HDF5

bool hdf5_available() { return true; }

namespace {

void put(H5::Group &g, const char *name, const std::vector<float> &v, int rank, const hsize_t *dims) {
    H5::DataSpace space(rank, dims);
    H5::DataSet ds = g.createDataSet(name, H5::PredType::NATIVE_FLOAT, space);
    ds.write(v.data(), H5::PredType::NATIVE_FLOAT);
}

// XDMF of one HDF5 file (hdf5Helper.py:58-139; without the force
// attribute for the time average, hdf5Helper.py:140-200).
void write_xmf(const std::string &xmfname, const std::string &h5name, int nz, int ny, int nx,
               double dx, bool with_force) {
    FILE *f = std::fopen(xmfname.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + xmfname);
    auto item = [&](const char *path) {
        std::fprintf(f, "<DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Format=\"HDF\">\n", nz, ny, nx);
        std::fprintf(f, "%s:%s\n", h5name.c_str(), path);
        std::fprintf(f, "</DataItem>\n");
    };
    std::fprintf(f, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
    std::fprintf(f, "<Xdmf xmlns:xi=\"http://www.w3.org/2003/XInclude\" Version=\"2.1\">\n<Domain>\n");
    std::fprintf(f, "<Grid Name=\"my_Grid\" GridType=\"Uniform\">\n");
    std::fprintf(f, "<Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%d %d %d\">\n</Topology>\n", nz, ny, nx);
    std::fprintf(f, "<Geometry GeometryType=\"Origin_DxDyDz\">\n");
    std::fprintf(f, "<DataItem Dimensions=\"3\" NumberType=\"Integer\" Format=\"XML\">\n0 0 0\n</DataItem>\n");
    std::fprintf(f, "<DataItem Dimensions=\"3\" NumberType=\"Integer\" Format=\"XML\">\n%g %g %g\n</DataItem>\n", dx, dx, dx);
    std::fprintf(f, "</Geometry>\n");
    std::fprintf(f, "<Attribute Name=\"velocity\" AttributeType=\"Vector\" Center=\"Node\">\n");
    std::fprintf(f, "<DataItem ItemType=\"Function\" Function=\"JOIN($0, $1, $2)\" Dimensions=\"%d %d %d 3\">\n", nz, ny, nx);
    item("/velo_group/x_velo");
    item("/velo_group/y_velo");
    item("/velo_group/z_velo");
    std::fprintf(f, "</DataItem>\n</Attribute>\n");
    if (with_force) {
        std::fprintf(f, "<Attribute Name=\"force\" AttributeType=\"Vector\" Center=\"Node\">\n");
        std::fprintf(f, "<DataItem ItemType=\"Function\" Function=\"JOIN($0, $1, $2)\" Dimensions=\"%d %d %d 3\">\n", nz, ny, nx);
        item("/force_group/x_force");
        item("/force_group/y_force");
        item("/force_group/z_force");
        std::fprintf(f, "</DataItem>\n</Attribute>\n");
    }
    std::fprintf(f, "<Attribute Name=\"pressure\" AttributeType=\"Scalar\" Center=\"Node\">\n");
    item("/pres_group/presmag");
    std::fprintf(f, "</Attribute>\n");
    std::fprintf(f, "<Attribute Name=\"velocityMagnitude\" AttributeType=\"Scalar\" Center=\"Node\">\n");
    item("/velo_group/velmag");
    std::fprintf(f, "</Attribute>\n</Grid>\n</Domain>\n</Xdmf>\n");
    std::fclose(f);
}

// Velocity in physical units, pressure relative to node pRefIdx and the
// speed of one field (processTLBM.py:82-104, 144-165): the fields are
// divided by `count` first (1 for a dump, the step count for time-average
// sums).
struct PhysicalFields {
    std::vector<float> u, v, w, pres, mag;
};

PhysicalFields physical(const Problem &p, const std::vector<double> &ux,
                        const std::vector<double> &uy, const std::vector<double> &uz,
                        const std::vector<double> &rho, double count) {
    const size_t n = ux.size();
    const double u_conv = p.tConvFact / p.lConvFact;  // processTLBM.py:51
    PhysicalFields r;
    r.u.resize(n);
    r.v.resize(n);
    r.w.resize(n);
    r.pres.resize(n);
    r.mag.resize(n);
    for (size_t i = 0; i < n; ++i) {
        r.u[i] = float(ux[i] / count / u_conv);
        r.v[i] = float(uy[i] / count / u_conv);
        r.w[i] = float(uz[i] / count / u_conv);
        r.pres[i] = float(rho[i] / count * p.pConvFact);
    }
    const float pref = (p.pRefIdx >= 0 && size_t(p.pRefIdx) < n) ? r.pres[size_t(p.pRefIdx)] : 0.f;
    for (size_t i = 0; i < n; ++i) {
        r.pres[i] -= pref;
        r.mag[i] = std::sqrt(r.u[i] * r.u[i] + r.v[i] * r.v[i] + r.w[i] * r.w[i]);
    }
    return r;
}

// velo_group {x_velo, y_velo, z_velo} (flat) and velmag, pres_group/presmag
// (shaped nz, ny, nx), hdf5Helper.py:4-56.
void put_velocity_pressure(H5::H5File &file, const Problem &p, const PhysicalFields &ph) {
    const hsize_t flat[1] = {hsize_t(ph.u.size())};
    const hsize_t grid[3] = {hsize_t(p.nz), hsize_t(p.ny), hsize_t(p.nx)};
    H5::Group velo = file.createGroup("velo_group");
    put(velo, "x_velo", ph.u, 1, flat);
    put(velo, "y_velo", ph.v, 1, flat);
    put(velo, "z_velo", ph.w, 1, flat);
    put(velo, "velmag", ph.mag, 3, grid);
}

double grid_dx(const Problem &p) { return p.nx > 1 ? p.lx_p / double(p.nx - 1) : p.lx_p; }

}  // namespace

void write_dump_h5(const std::string &h5name, const std::string &xmfname, const Problem &p,
                   const std::vector<double> &ux, const std::vector<double> &uy,
                   const std::vector<double> &uz, const std::vector<double> &rho,
                   const std::vector<double> &fx, const std::vector<double> &fy,
                   const std::vector<double> &fz) {
    const size_t n = ux.size();
    const double u_conv = p.tConvFact / p.lConvFact;
    const double f_conv = p.lConvFact / (u_conv * u_conv);  // processTLBM.py:106
    const PhysicalFields ph = physical(p, ux, uy, uz, rho, 1.0);
    std::vector<float> Fx(n), Fy(n), Fz(n);
    for (size_t i = 0; i < n; ++i) {
        Fx[i] = float(fx[i] * f_conv);
        Fy[i] = float(fy[i] * f_conv);
        Fz[i] = float(fz[i] * f_conv);
    }
    H5::H5File file(h5name, H5F_ACC_TRUNC);
    const hsize_t flat[1] = {hsize_t(n)};
    const hsize_t grid[3] = {hsize_t(p.nz), hsize_t(p.ny), hsize_t(p.nx)};
    put_velocity_pressure(file, p, ph);
    H5::Group force = file.createGroup("force_group");
    put(force, "x_force", Fx, 1, flat);
    put(force, "y_force", Fy, 1, flat);
    put(force, "z_force", Fz, 1, flat);
    H5::Group pg = file.createGroup("pres_group");
    put(pg, "presmag", ph.pres, 3, grid);
    write_xmf(xmfname, h5name, p.nz, p.ny, p.nx, grid_dx(p), true);
}

void write_time_avg_h5(const std::string &h5name, const std::string &xmfname, const Problem &p,
                       const std::vector<double> &u_sum, const std::vector<double> &v_sum,
                       const std::vector<double> &w_sum, const std::vector<double> &rho_sum) {
    // the sums divided by Num_ts (processTLBM.py:144-147), then as a dump
    const PhysicalFields ph = physical(p, u_sum, v_sum, w_sum, rho_sum, double(p.numTs));
    H5::H5File file(h5name, H5F_ACC_TRUNC);
    const hsize_t grid[3] = {hsize_t(p.nz), hsize_t(p.ny), hsize_t(p.nx)};
    put_velocity_pressure(file, p, ph);
    H5::Group pg = file.createGroup("pres_group");
    put(pg, "presmag", ph.pres, 3, grid);
    write_xmf(xmfname, h5name, p.nz, p.ny, p.nx, grid_dx(p), false);
}

void read_restart(const std::string &path, size_t n, std::vector<double> &ux,
                  std::vector<double> &uy, std::vector<double> &uz, std::vector<double> &rho) {
    H5::H5File file(path, H5F_ACC_RDONLY);
    const char *names[4] = {"velocity/x", "velocity/y", "velocity/z", "density/rho"};
    std::vector<double> *dst[4] = {&ux, &uy, &uz, &rho};
    std::vector<float> tmp;
    for (int k = 0; k < 4; ++k) {
        H5::DataSet ds = file.openDataSet(names[k]);
        const hssize_t count = ds.getSpace().getSimpleExtentNpoints();
        if (count != hssize_t(n))
            throw std::runtime_error(path + ": " + names[k] + " has " + std::to_string(count) +
                                     " values, expected " + std::to_string(n));
        tmp.resize(n);
        ds.read(tmp.data(), H5::PredType::NATIVE_FLOAT);
        dst[k]->assign(tmp.begin(), tmp.end());
    }
}

#else

bool hdf5_available() { return false; }

void write_dump_h5(const std::string &, const std::string &, const Problem &,
                   const std::vector<double> &, const std::vector<double> &,
                   const std::vector<double> &, const std::vector<double> &,
                   const std::vector<double> &, const std::vector<double> &,
                   const std::vector<double> &) {
    throw std::runtime_error("built without HDF5: out<k>.h5 cannot be written");
}

void read_restart(const std::string &, size_t, std::vector<double> &, std::vector<double> &,
                  std::vector<double> &, std::vector<double> &) {
    throw std::runtime_error("built without HDF5: restart.h5 cannot be read");
}

void write_time_avg_h5(const std::string &, const std::string &, const Problem &,
                       const std::vector<double> &, const std::vector<double> &,
                       const std::vector<double> &, const std::vector<double> &) {
    throw std::runtime_error("built without HDF5: timeAvg.h5 cannot be written");
}

#endif

}  // namespace tlbm_host
