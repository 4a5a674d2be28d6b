// tlbm_restart.cpp -- restart.h5 from dump k of a run, as
// python/tLBM_create_restart.py:34-68 makes it: density<k>, ux<k>, uy<k>,
// uz<k>.b_dat reordered to global order through ordering.b_dat and stored
// as float32 datasets /velocity/{x,y,z} and /density/rho.
//
//   tlbm_restart <dumpNum> [f32|f64]   (value type of the .b_dat files)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef TLBM_WITH_HDF5
#include "H5Cpp.h"
#endif

namespace {

template <class T>
std::vector<T> read_all(const std::string &name) {
    FILE *fp = std::fopen(name.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open " + name);
    std::fseek(fp, 0, SEEK_END);
    const long bytes = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::vector<T> v(size_t(bytes) / sizeof(T));
    const size_t got = std::fread(v.data(), sizeof(T), v.size(), fp);
    std::fclose(fp);
    if (got != v.size()) throw std::runtime_error("short read of " + name);
    return v;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: tlbm_restart <dumpNum> [f32|f64]\n");
        return 2;
    }
#ifndef TLBM_WITH_HDF5
    std::fprintf(stderr, "tlbm_restart: built without HDF5\n");
    return 1;
#else
    try {
        const std::string k = argv[1];
        const bool f64 = argc > 2 && std::strcmp(argv[2], "f64") == 0;
        const std::vector<int32_t> order = read_all<int32_t>("ordering.b_dat");
        const size_t n = order.size();
        const char *fields[4] = {"ux", "uy", "uz", "density"};
        std::vector<float> out[4];
        for (int i = 0; i < 4; ++i) {
            const std::string fn = std::string(fields[i]) + k + ".b_dat";
            std::vector<double> raw;
            if (f64) {
                raw = read_all<double>(fn);
            } else {
                const std::vector<float> r = read_all<float>(fn);
                raw.assign(r.begin(), r.end());
            }
            if (raw.size() != n) throw std::runtime_error(fn + " does not match ordering.b_dat");
            out[i].assign(n, 0.f);
            for (size_t j = 0; j < n; ++j) out[i][size_t(order[j])] = float(raw[j]);
        }
        H5::H5File file("restart.h5", H5F_ACC_EXCL);
        const hsize_t dims[1] = {hsize_t(n)};
        H5::DataSpace space(1, dims);
        H5::Group velo = file.createGroup("velocity");
        const char *vn[3] = {"x", "y", "z"};
        for (int i = 0; i < 3; ++i)
            velo.createDataSet(vn[i], H5::PredType::NATIVE_FLOAT, space)
                .write(out[i].data(), H5::PredType::NATIVE_FLOAT);
        H5::Group dens = file.createGroup("density");
        dens.createDataSet("rho", H5::PredType::NATIVE_FLOAT, space)
            .write(out[3].data(), H5::PredType::NATIVE_FLOAT);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tlbm_restart: %s\n", e.what());
        return 1;
    } catch (const H5::Exception &e) {
        std::fprintf(stderr, "tlbm_restart: HDF5: %s\n", e.getDetailMsg().c_str());
        return 1;
    }
    return 0;
#endif
}
