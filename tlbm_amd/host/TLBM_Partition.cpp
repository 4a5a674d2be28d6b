// TLBM_Partition.cpp -- the partition object over the GPU C-ABI.
#include "TLBM_Partition.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "h5_io.h"
#include "lattice.h"

namespace {

void check_rc(int rc, const char *what) {
    if (rc != TLBM_OK)
        throw std::runtime_error(std::string(what) + ": " + tlbm_gpu_last_error());
}

int storage_from_env() {
    const char *s = std::getenv("TLBM_STORAGE");
    if (!s || !*s || std::strcmp(s, "f32") == 0) return TLBM_STORAGE_F32;
    if (std::strcmp(s, "f64") == 0) return TLBM_STORAGE_F64;
    if (std::strcmp(s, "f16") == 0) return TLBM_STORAGE_F16;
    if (std::strcmp(s, "f32x") == 0) return TLBM_STORAGE_F32X;
    throw std::invalid_argument(std::string("TLBM_STORAGE must be f64, f32, f16 or f32x, not ") + s);
}

bool env_flag(const char *name) {
    const char *s = std::getenv(name);
    return s && *s && std::strcmp(s, "0") != 0;
}

// Whole-file integer reader for ndType.lbm / parts.lbm (one int per line,
// tLBM_write_params.py:103-108).  Reads in blocks; fails on a non-integer.
template <class F>
size_t for_each_int(const char *path, F &&fn) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) throw std::runtime_error(std::string("Could not open ") + path);
    std::vector<char> buf(size_t(1) << 22);
    size_t count = 0, carry = 0;
    bool eof = false;
    while (!eof) {
        const size_t got = std::fread(buf.data() + carry, 1, buf.size() - carry - 1, fp);
        eof = got == 0 || std::feof(fp);
        size_t len = carry + got;
        size_t end = len;
        if (!eof) {  // keep a partial number for the next block
            while (end > 0 && buf[end - 1] != ' ' && buf[end - 1] != '\n' && buf[end - 1] != '\r' &&
                   buf[end - 1] != '\t')
                --end;
        }
        buf[end] = '\0';
        char *p = buf.data();
        char *stop = buf.data() + end;
        while (p < stop) {
            while (p < stop && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
            if (p >= stop) break;
            char *q = nullptr;
            errno = 0;
            const long v = std::strtol(p, &q, 10);
            if (q == p || errno != 0) {
                std::fclose(fp);
                throw std::runtime_error(std::string(path) + ": malformed entry " +
                                         std::to_string(count));
            }
            fn(count++, v);
            p = q;
        }
        carry = len - end;
        std::memmove(buf.data(), buf.data() + end, carry);
        if (eof && carry) throw std::runtime_error(std::string(path) + ": truncated entry");
    }
    std::fclose(fp);
    return count;
}

}  // namespace

TLBM_Partition::TLBM_Partition(int rank, int size) : rank_(rank) {
    if (size != 1 || rank != 0)
        throw std::invalid_argument("one GPU holds the whole lattice: rank 0 of size 1 only");
    thisProblem.load_input();
    const std::string &lt = thisProblem.latticeType;
    if (lt == "D3Q15") {
        Q_ = 15;
        ex_.assign(tlbm::D3Q15::ex, tlbm::D3Q15::ex + 15);
        ey_.assign(tlbm::D3Q15::ey, tlbm::D3Q15::ey + 15);
        ez_.assign(tlbm::D3Q15::ez, tlbm::D3Q15::ez + 15);
    } else if (lt == "D3Q19") {
        Q_ = 19;
        ex_.assign(tlbm::D3Q19::ex, tlbm::D3Q19::ex + 19);
        ey_.assign(tlbm::D3Q19::ey, tlbm::D3Q19::ey + 19);
        ez_.assign(tlbm::D3Q19::ez, tlbm::D3Q19::ez + 19);
    } else if (lt == "D3Q27") {
        Q_ = 27;
        ex_.assign(tlbm::D3Q27::ex, tlbm::D3Q27::ex + 27);
        ey_.assign(tlbm::D3Q27::ey, tlbm::D3Q27::ey + 27);
        ez_.assign(tlbm::D3Q27::ez, tlbm::D3Q27::ez + 27);
    } else {
        throw std::invalid_argument("Invalid Lattice Structure!");
    }
    const Problem &p = thisProblem;
    if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0)
        throw std::invalid_argument("params.lbm: nx, ny, nz must be positive");
    if (p.plotFreq <= 0 || p.tsRepFreq <= 0)
        throw std::invalid_argument("params.lbm: plotFreq and tsRepFreq must be positive");
    storage_ = storage_from_env();
    std::vector<uint8_t> nd;
    load_ndType(nd);
    check_parts();
    tlbm_params gp{};
    gp.lattice = Q_;
    gp.dynamics = p.dynamics;
    gp.nx = p.nx;
    gp.ny = p.ny;
    gp.nz = p.nz;
    gp.storage = storage_;
    gp.exact_weights = env_flag("TLBM_EXACT_WEIGHTS") ? 1 : 0;
    gp.reserved = 0;
    gp.cs = p.cs;
    gp.rho_lbm = p.rhoLBM;
    gp.u_lbm = p.uLBM;
    gp.omega = p.omega;
    const int rc = tlbm_gpu_create(&gp, nd.data(), &gpu_);
    if (rc == TLBM_EINVAL) throw std::invalid_argument(tlbm_gpu_last_error());
    check_rc(rc, "tlbm_gpu_create");
    if (p.restartFlag) load_restart_data();
    if (p.timeAvgFlag == 1) check_rc(tlbm_gpu_time_average(gpu_, 1), "time average");
    write_node_ordering();
}

TLBM_Partition::~TLBM_Partition() { tlbm_gpu_destroy(gpu_); }

namespace {

bool valid_type(long v) { return v == 0 || v == 1 || v == 2 || v == 3 || v == 5; }

// CRC-32 (IEEE 802.3, reflected, polynomial 0xEDB88320) of a whole file
// and its size; false when the file cannot be opened.
bool file_crc32(const char *path, uint64_t &size, uint32_t &crc) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        ready = true;
    }
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return false;
    std::vector<unsigned char> buf(size_t(1) << 22);
    uint32_t c = 0xFFFFFFFFu;
    size = 0;
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), fp)) > 0) {
        for (size_t i = 0; i < got; ++i) c = table[(c ^ buf[i]) & 0xFFu] ^ (c >> 8);
        size += got;
    }
    std::fclose(fp);
    crc = c ^ 0xFFFFFFFFu;
    return true;
}

uint64_t le64(const unsigned char *b) {
    uint64_t v = 0;
    for (int k = 7; k >= 0; --k) v = (v << 8) | b[k];
    return v;
}

// ndType.bin: "TLBMNDT2", then the node count, the size and the CRC-32 of
// the ndType.lbm it was written from (zeros when none), each a
// little-endian uint64, then one byte per node in global order.  Used in
// place of ndType.lbm when that file is absent, or when it is the very file
// recorded (same size and CRC-32); otherwise the text file is read.
bool read_ndtype_sidecar(std::vector<uint8_t> &nd, size_t N) {
    struct stat sb;
    if (stat("ndType.bin", &sb) != 0) return false;
    FILE *fp = std::fopen("ndType.bin", "rb");
    if (!fp) throw std::runtime_error("Could not open ndType.bin");
    unsigned char head[32];
    if (std::fread(head, 1, 32, fp) != 32 || std::memcmp(head, "TLBMNDT2", 8) != 0) {
        std::fclose(fp);
        throw std::runtime_error("ndType.bin: not an ndType sidecar");
    }
    const uint64_t n = le64(head + 8), text_size = le64(head + 16), text_crc = le64(head + 24);
    struct stat st;
    if (stat("ndType.lbm", &st) == 0) {
        uint64_t size = 0;
        uint32_t crc = 0;
        const bool same = uint64_t(st.st_size) == text_size && file_crc32("ndType.lbm", size, crc) &&
                          size == text_size && uint64_t(crc) == text_crc;
        if (!same) {
            std::fclose(fp);
            std::fprintf(stderr, "ndType.bin was not written from this ndType.lbm; reading ndType.lbm\n");
            return false;
        }
    }
    if (n != N) {
        std::fclose(fp);
        throw std::runtime_error("ndType.bin has " + std::to_string(n) + " entries, expected " +
                                 std::to_string(N));
    }
    nd.assign(N, 0);
    const size_t got = std::fread(nd.data(), 1, N, fp);
    const bool extra = std::fgetc(fp) != EOF;
    std::fclose(fp);
    if (got != N || extra)
        throw std::runtime_error("ndType.bin: payload is not " + std::to_string(N) + " bytes");
    for (size_t i = 0; i < N; ++i)
        if (!valid_type(nd[i]))
            throw std::invalid_argument("ndType.bin: node " + std::to_string(i) + " has type " +
                                        std::to_string(int(nd[i])) + "; valid types are 0,1,2,3,5");
    return true;
}

}  // namespace

void TLBM_Partition::load_ndType(std::vector<uint8_t> &nd) {
    const size_t N = size_t(thisProblem.nx) * thisProblem.ny * thisProblem.nz;
    if (read_ndtype_sidecar(nd, N)) return;
    nd.assign(N, 0);
    const size_t n = for_each_int("ndType.lbm", [&](size_t i, long v) {
        if (i >= N) throw std::runtime_error("ndType.lbm has more than nx*ny*nz entries");
        if (!valid_type(v))
            throw std::invalid_argument("ndType.lbm: node " + std::to_string(i) + " has type " +
                                        std::to_string(v) + "; valid types are 0,1,2,3,5");
        nd[i] = uint8_t(v);
    });
    if (n != N)
        throw std::runtime_error("ndType.lbm has " + std::to_string(n) + " entries, expected " +
                                 std::to_string(N));
}

void TLBM_Partition::check_parts() {
    // partition ids are read for format compatibility; one GPU takes all nodes
    std::ifstream probe("parts.lbm");
    if (!probe.is_open()) return;
    probe.close();
    const size_t N = size_t(thisProblem.nx) * thisProblem.ny * thisProblem.nz;
    const size_t n = for_each_int("parts.lbm", [&](size_t, long v) {
        if (v < 0) throw std::runtime_error("parts.lbm: negative partition id");
    });
    if (n != N)
        throw std::runtime_error("parts.lbm has " + std::to_string(n) + " entries, expected " +
                                 std::to_string(N));
}

void TLBM_Partition::write_node_ordering() {
    // global node number of each dumped entry (TLBM_Partition.cpp:542-562):
    // identity on one partition
    const int N = get_num_global_nodes();
    std::vector<int32_t> order(static_cast<size_t>(N));
    for (int i = 0; i < N; ++i) order[size_t(i)] = i;
    FILE *fp = std::fopen("ordering.b_dat", "wb");
    if (!fp) throw std::runtime_error("Error opening file to write node ordering");
    const size_t w = std::fwrite(order.data(), sizeof(int32_t), order.size(), fp);
    std::fclose(fp);
    if (w != order.size()) throw std::runtime_error("short write of ordering.b_dat");
}

void TLBM_Partition::load_restart_data() {
    if (rank_ == 0) std::printf("Loading Restart Data \n");
    const size_t N = size_t(get_num_global_nodes());
    std::vector<double> ux, uy, uz, rho;
    tlbm_host::read_restart("restart.h5", N, ux, uy, uz, rho);
    check_rc(tlbm_gpu_init_from_macros(gpu_, ux.data(), uy.data(), uz.data(), rho.data()),
             "restart");
    pop_cached_ = false;
}

int TLBM_Partition::get_num_ts() { return thisProblem.numTs; }
int TLBM_Partition::get_ts_rep_freq() { return thisProblem.tsRepFreq; }
int TLBM_Partition::get_plot_freq() { return thisProblem.plotFreq; }
int TLBM_Partition::get_warmupTs() { return thisProblem.warmupTs; }
int TLBM_Partition::is_restart() { return thisProblem.restartFlag; }
int TLBM_Partition::get_time_avg_flag() { return thisProblem.timeAvgFlag; }
int TLBM_Partition::get_cut_size() { return 0; }

int TLBM_Partition::get_num_global_nodes() {
    return thisProblem.nx * thisProblem.ny * thisProblem.nz;
}

// TLBM_Partition.cpp:716-728
LatticeIndex TLBM_Partition::get_xyz_index(int gInd) {
    const int nx = thisProblem.nx, ny = thisProblem.ny;
    LatticeIndex r;
    r.Z = gInd / (nx * ny);
    r.Y = (gInd - r.Z * nx * ny) / nx;
    r.X = gInd - r.Z * nx * ny - r.Y * nx;
    return r;
}

// TLBM_Partition.cpp:691-712: periodic in x, y and z
int TLBM_Partition::get_tgt_index(int gInd, int ex, int ey, int ez) {
    const LatticeIndex s = get_xyz_index(gInd);
    LatticeIndex t;
    t.X = (s.X + ex) % thisProblem.nx;
    t.Y = (s.Y + ey) % thisProblem.ny;
    t.Z = (s.Z + ez) % thisProblem.nz;
    if (t.X < 0) t.X += thisProblem.nx;
    if (t.Y < 0) t.Y += thisProblem.ny;
    if (t.Z < 0) t.Z += thisProblem.nz;
    return get_gInd(t);
}

int TLBM_Partition::get_gInd(int x, int y, int z) {
    return x + y * thisProblem.nx + z * thisProblem.nx * thisProblem.ny;
}

int TLBM_Partition::get_gInd(LatticeIndex i) { return get_gInd(i.X, i.Y, i.Z); }

void TLBM_Partition::print_adjacency(const int nd) {
    std::cout << "adjacency for nd " << nd << ": ";
    for (int s = 0; s < Q_; ++s) std::cout << get_tgt_index(nd, ex_[s], ey_[s], ez_[s]) << ", ";
    std::cout << std::endl;
}

double TLBM_Partition::get_data_member(int nd, int spd) {
    if (nd < 0 || nd >= get_num_global_nodes() || spd < 0 || spd >= Q_)
        throw std::out_of_range("get_data_member: node or speed out of range");
    if (!pop_cached_) {
        pop_cache_.resize(size_t(get_num_global_nodes()) * Q_);
        check_rc(tlbm_gpu_get_populations(gpu_, pop_cache_.data()), "get populations");
        pop_cached_ = true;
    }
    return pop_cache_[size_t(nd) * Q_ + spd];
}

void TLBM_Partition::set_data_member(double val, int nd, int spd) {
    (void)get_data_member(nd, spd);  // range check and cache fill
    pop_cache_[size_t(nd) * Q_ + spd] = val;
    pop_dirty_ = true;
}

void TLBM_Partition::copy_populations(double *f) {
    sync_populations_to_device();
    check_rc(tlbm_gpu_get_populations(gpu_, f), "get populations");
}

void TLBM_Partition::load_populations(const double *f) {
    check_rc(tlbm_gpu_set_populations(gpu_, f), "set populations");
    pop_cached_ = false;
    pop_dirty_ = false;
}

void TLBM_Partition::process_node_list(double *, double *, const std::set<int> &) {
    throw std::logic_error(
        "process_node_list: the GPU step updates every node at once; call take_LBM_time_step");
}

void TLBM_Partition::process_node_list(float *, float *, const std::set<int> &) {
    throw std::logic_error(
        "process_node_list: the GPU step updates every node at once; call take_LBM_time_step");
}

void TLBM_Partition::sync_populations_to_device() {
    if (pop_dirty_) {
        check_rc(tlbm_gpu_set_populations(gpu_, pop_cache_.data()), "set populations");
        pop_dirty_ = false;
    }
}

void TLBM_Partition::take_LBM_time_step(bool isEven) {
    (void)isEven;
    sync_populations_to_device();
    check_rc(tlbm_gpu_step(gpu_, 1), "time step");
    pop_cached_ = false;
}

void TLBM_Partition::synchronize() { check_rc(tlbm_gpu_sync(gpu_), "sync"); }

void TLBM_Partition::check_finite(int ts) {
    sync_populations_to_device();
    double red[5];
    const int rc = tlbm_gpu_reduce(gpu_, red);
    if (rc == TLBM_ENONFINITE)
        throw std::runtime_error("time step " + std::to_string(ts) + ": " + tlbm_gpu_last_error());
    check_rc(rc, "state check");
}

template <class Real>
void TLBM_Partition::write_field(const std::string &name, const std::vector<double> &v) {
    std::vector<Real> out(v.begin(), v.end());
    FILE *fp = std::fopen(name.c_str(), "wb");
    if (!fp) throw std::runtime_error("Error opening file to write " + name);
    const size_t w = std::fwrite(out.data(), sizeof(Real), out.size(), fp);
    std::fclose(fp);
    if (w != out.size()) throw std::runtime_error("short write of " + name);
}

void TLBM_Partition::write_data() {
    sync_populations_to_device();
    const size_t N = size_t(get_num_global_nodes());
    std::vector<double> ux(N), uy(N), uz(N), rho(N), fx(N), fy(N), fz(N);
    check_rc(tlbm_gpu_get_macros(gpu_, ux.data(), uy.data(), uz.data(), rho.data()), "moments");
    check_rc(tlbm_gpu_get_forces(gpu_, fx.data(), fy.data(), fz.data()), "forces");
    const std::string k = std::to_string(dataWriteNum);
    const std::vector<double> *fields[7] = {&ux, &uy, &uz, &rho, &fx, &fy, &fz};
    const char *names[7] = {"ux", "uy", "uz", "density", "Fx", "Fy", "Fz"};
    for (int i = 0; i < 7; ++i) {
        const std::string fn = std::string(names[i]) + k + ".b_dat";
        if (storage_ == TLBM_STORAGE_F64)
            write_field<double>(fn, *fields[i]);
        else
            write_field<float>(fn, *fields[i]);
    }
    if (env_flag("TLBM_HDF5"))
        tlbm_host::write_dump_h5("out" + k + ".h5", "data" + k + ".xmf", thisProblem, ux, uy, uz, rho,
                                 fx, fy, fz);
    ++dataWriteNum;
}

void TLBM_Partition::write_time_avg_data() {
    const size_t N = size_t(get_num_global_nodes());
    std::vector<double> u(N), v(N), w(N), r(N);
    int64_t n = 0;
    check_rc(tlbm_gpu_get_time_average(gpu_, u.data(), v.data(), w.data(), r.data(), &n),
             "time average");
    const std::vector<double> *fields[4] = {&u, &v, &w, &r};
    const char *names[4] = {"uAvg.b_dat", "vAvg.b_dat", "wAvg.b_dat", "rhoAvg.b_dat"};
    for (int i = 0; i < 4; ++i) {
        if (storage_ == TLBM_STORAGE_F64)
            write_field<double>(names[i], *fields[i]);
        else
            write_field<float>(names[i], *fields[i]);
    }
    if (env_flag("TLBM_HDF5"))
        tlbm_host::write_time_avg_h5("timeAvg.h5", "timeAvg.xmf", thisProblem, u, v, w, r);
}
