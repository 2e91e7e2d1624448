// io.cpp -- host-side writers of the reference's output layout:
//   <root>/JET_VTK/Re<Re>_We<We>/case_info.txt            meanF/meanFieldsRT.cuh:93-118
//   .../mean_profiles/tke_{total,avg,diff}.bin (float32)    :73-91, :366-375
//   .../mean_profiles/radial_sum_{uy,uy2,ur,ur2,uruy}.bin,
//       radial_count.bin, radial_profile_metadata.txt        :229-364
//   .../vti_data/lbm_%08d.vti (VTK XML ImageData, base64)   io/save_data.cuh:22-215
// The byte layouts are the reference's so that its post-processing reads
// them unchanged; radial_sum_phi{,2}.bin are additional files.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "../../include/mcg.h"

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const std::string &m) {
    g_io_err = m;
    return code;
}

std::string case_folder(const mcg_case &c) {
    char buf[96];
    std::snprintf(buf, sizeof(buf), "Re%d_We%d", int(std::lround(c.Re)), int(std::lround(c.We)));
    return buf;
}

// RFC 4648 base64 with '=' padding
void b64_append(std::string &out, const unsigned char *d, size_t n) {
    static const char T[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    out.reserve(out.size() + (n + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 3 <= n; i += 3) {
        const uint32_t v = (uint32_t(d[i]) << 16) | (uint32_t(d[i + 1]) << 8) | d[i + 2];
        out.push_back(T[v >> 18]);
        out.push_back(T[(v >> 12) & 63]);
        out.push_back(T[(v >> 6) & 63]);
        out.push_back(T[v & 63]);
    }
    if (i < n) {
        const uint32_t v = (uint32_t(d[i]) << 16) | (i + 1 < n ? uint32_t(d[i + 1]) << 8 : 0u);
        out.push_back(T[v >> 18]);
        out.push_back(T[(v >> 12) & 63]);
        out.push_back(i + 1 < n ? T[(v >> 6) & 63] : '=');
        out.push_back('=');
    }
}

// [u64 byte count][payload] encoded as one base64 block (VTK binary, header_type UInt64)
std::string vtk_block(const std::vector<unsigned char> &payload) {
    std::vector<unsigned char> buf(8 + payload.size());
    const uint64_t nb = payload.size();
    std::memcpy(buf.data(), &nb, 8);
    if (!payload.empty()) std::memcpy(buf.data() + 8, payload.data(), payload.size());
    std::string s;
    b64_append(s, buf.data(), buf.size());
    return s;
}

bool write_file(const std::string &path, const void *data, size_t esize, size_t count) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t w = std::fwrite(data, esize, count, f);
    std::fclose(f);
    return w == count;
}

}  // namespace

extern "C" {

const char *mcg_io_last_error(void) { return g_io_err.c_str(); }

int mcg_output_dirs(const mcg_case *c, const char *root, char *case_dir, char *mean_dir, char *vti_dir, size_t cap) {
    if (!c || !root) return io_fail(MCG_E_INVALID, "null argument");
    namespace fs = std::filesystem;
    const fs::path cd = fs::path(root) / "JET_VTK" / case_folder(*c);
    const fs::path md = cd / "mean_profiles";
    const fs::path vd = cd / "vti_data";
    std::error_code ec;
    fs::create_directories(md, ec);
    if (ec) return io_fail(MCG_E_IO, "cannot create " + md.string());
    fs::create_directories(vd, ec);
    if (ec) return io_fail(MCG_E_IO, "cannot create " + vd.string());
    const std::string s[3] = {cd.string(), md.string(), vd.string()};
    char *dst[3] = {case_dir, mean_dir, vti_dir};
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        if (s[k].size() + 1 > cap) return io_fail(MCG_E_INVALID, "path buffer too small");
        std::memcpy(dst[k], s[k].c_str(), s[k].size() + 1);
    }
    return MCG_OK;
}

int mcg_write_case_info(const mcg_case *c, const char *case_dir) {
    if (!c || !case_dir) return io_fail(MCG_E_INVALID, "null argument");
    const std::string p = std::string(case_dir) + "/case_info.txt";
    FILE *f = std::fopen(p.c_str(), "w");
    if (!f) return io_fail(MCG_E_IO, "cannot write " + p);
    std::fprintf(f, "Re %d\n", int(std::lround(c->Re)));
    std::fprintf(f, "We %d\n", int(std::lround(c->We)));
    std::fprintf(f, "NX %u\n", unsigned(c->NX));
    std::fprintf(f, "NY %u\n", unsigned(c->NY));
    std::fprintf(f, "NZ %u\n", unsigned(c->NZ));
    std::fprintf(f, "NSTEP %d\n", c->NSTEP);
    std::fprintf(f, "NOUTPUT %d\n", c->NOUTPUT);
    std::fprintf(f, "NSTATS_START_STEP %d\n", c->NSTATS_START_STEP);
    std::fprintf(f, "NR_BINS %u\n", unsigned(c->NX / 2));
    std::fprintf(f, "write_vti_output %d\n", c->write_vti_output ? 1 : 0);
    std::fprintf(f, "jet_radius %.9g\n", double(float(c->jet_radius)));
    std::fprintf(f, "jet_diameter %.9g\n", double(2.0f * float(c->jet_radius)));
    std::fprintf(f, "jet_velocity %.9g\n", double(float(c->jet_velocity)));
    std::fclose(f);
    return MCG_OK;
}

int mcg_write_vti_arrays(const char *path, int32_t NX, int32_t NY, int32_t NZ, const float *rhor, const float *rhob,
                         const float *ux, const float *uy, const float *uz) {
    if (!path || !rhor || !rhob || !ux || !uy || !uz || NX <= 0 || NY <= 0 || NZ <= 0)
        return io_fail(MCG_E_INVALID, "bad arguments");
    const size_t n = size_t(NX) * NY * NZ;
    std::vector<unsigned char> pl(n * sizeof(float));
    auto block_of = [&](const float *a) {
        std::memcpy(pl.data(), a, n * sizeof(float));
        return vtk_block(pl);
    };
    const std::string e_rr = block_of(rhor);
    const std::string e_rb = block_of(rhob);
    {
        float *q = reinterpret_cast<float *>(pl.data());
        for (size_t i = 0; i < n; ++i) q[i] = rhor[i] + rhob[i];
    }
    const std::string e_rho = vtk_block(pl);
    std::vector<unsigned char> pu(3 * n * sizeof(float));
    {
        float *q = reinterpret_cast<float *>(pu.data());
        for (size_t i = 0; i < n; ++i) {
            q[3 * i] = ux[i];
            q[3 * i + 1] = uy[i];
            q[3 * i + 2] = uz[i];
        }
    }
    const std::string e_u = vtk_block(pu);
    FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(MCG_E_IO, std::string("cannot open ") + path);
    std::fprintf(f, "<?xml version=\"1.0\"?>\n");
    std::fprintf(f, "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
    std::fprintf(f, "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n", NX - 1, NY - 1, NZ - 1);
    std::fprintf(f, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", NX - 1, NY - 1, NZ - 1);
    std::fprintf(f, "      <PointData Scalars=\"rhor\" Vectors=\"u\">\n");
    const char *names[3] = {"rhor", "rhob", "rho"};
    const std::string *blocks[3] = {&e_rr, &e_rb, &e_rho};
    for (int k = 0; k < 3; ++k) {
        std::fprintf(f, "        <DataArray type=\"Float32\" Name=\"%s\" format=\"binary\">\n", names[k]);
        std::fwrite(blocks[k]->data(), 1, blocks[k]->size(), f);
        std::fprintf(f, "\n        </DataArray>\n");
    }
    std::fprintf(f, "        <DataArray type=\"Float32\" Name=\"u\" NumberOfComponents=\"3\" format=\"binary\">\n");
    std::fwrite(e_u.data(), 1, e_u.size(), f);
    std::fprintf(f, "\n        </DataArray>\n");
    std::fprintf(f, "      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n");
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? MCG_OK : io_fail(MCG_E_IO, std::string("write error on ") + path);
}

int mcg_write_radial_files(const mcg_case *c, const char *mean_dir, const double *sum_uy, const double *sum_uy2,
                           const double *sum_ur, const double *sum_ur2, const double *sum_uruy, const uint64_t *count,
                           const double *sum_phi, const double *sum_phi2, int32_t started, int64_t start_step,
                           int64_t samples) {
    if (!c || !mean_dir) return io_fail(MCG_E_INVALID, "null argument");
    const std::string d(mean_dir);
    const size_t nb = size_t(c->NY) * size_t(c->NX / 2);
    bool ok = true;
    // the bins are written only once statistics have samples (meanFieldsRT.cuh:299-303)
    if (started && samples > 0) {
        const double *arr[5] = {sum_uy, sum_uy2, sum_ur, sum_ur2, sum_uruy};
        const char *nm[5] = {"radial_sum_uy", "radial_sum_uy2", "radial_sum_ur", "radial_sum_ur2", "radial_sum_uruy"};
        for (int k = 0; k < 5; ++k)
            if (arr[k]) ok = write_file(d + "/" + nm[k] + ".bin", arr[k], sizeof(double), nb) && ok;
        if (count) ok = write_file(d + "/radial_count.bin", count, sizeof(uint64_t), nb) && ok;
        if (sum_phi) ok = write_file(d + "/radial_sum_phi.bin", sum_phi, sizeof(double), nb) && ok;
        if (sum_phi2) ok = write_file(d + "/radial_sum_phi2.bin", sum_phi2, sizeof(double), nb) && ok;
    }
    const std::string p = d + "/radial_profile_metadata.txt";
    FILE *f = std::fopen(p.c_str(), "w");
    if (!f) return io_fail(MCG_E_IO, "cannot write " + p);
    const int every = c->NSTATS_SAMPLE > 0 ? c->NSTATS_SAMPLE : 1;
    std::fprintf(f, "NX %u\n", unsigned(c->NX));
    std::fprintf(f, "NY %u\n", unsigned(c->NY));
    std::fprintf(f, "NZ %u\n", unsigned(c->NZ));
    std::fprintf(f, "NR_BINS %u\n", unsigned(c->NX / 2));
    std::fprintf(f, "NradialProfileCells %u\n", unsigned(nb));
    std::fprintf(f, "NOUTPUT %d\n", c->NOUTPUT);
    std::fprintf(f, "NPROFILE_OUTPUT %d\n", c->NOUTPUT);
    std::fprintf(f, "NSTATS_SAMPLE %d\n", every);
    std::fprintf(f, "NSTATS_START_STEP %d\n", c->NSTATS_START_STEP);
    std::fprintf(f, "Re %d\n", int(std::lround(c->Re)));
    std::fprintf(f, "We %d\n", int(std::lround(c->We)));
    std::fprintf(f, "D %.9g\n", double(2.0f * float(c->jet_radius)));
    std::fprintf(f, "U_jet %.9g\n", double(float(c->jet_velocity)));
    std::fprintf(f, "jet_radius %.9g\n", double(float(c->jet_radius)));
    std::fprintf(f, "jet_velocity %.9g\n", double(float(c->jet_velocity)));
    std::fprintf(f, "uy_avg_started %d\n", started ? 1 : 0);
    std::fprintf(f, "start_step %u\n", unsigned(start_step));
    std::fprintf(f, "uy_avg_start_step %u\n", unsigned(start_step));
    std::fprintf(f, "uy_avg_samples %u\n", unsigned(samples));
    std::fprintf(f, "radial_sample_count %u\n", unsigned(samples));
    if (started && samples > 0) {
        std::fprintf(f, "uy_avg_first_sample_step %u\n", unsigned(start_step));
        std::fprintf(f, "uy_avg_last_sample_step %u\n", unsigned(start_step + (samples - 1) * every));
    }
    std::fclose(f);
    if (!ok) return io_fail(MCG_E_IO, "could not write one or more radial files in " + d);
    return MCG_OK;
}

// Running TKE average of update_tke_average (kernels.cu:154-166) and the
// delta bookkeeping of update_tke_state (meanF/meanFieldsRT.cuh:191-227), FP32
// as in the reference.  state = {tke_avg, tke_avg_prev, first_flag}.
int mcg_tke_update(float state[3], float tke_total, int64_t step, int32_t noutput, float out[3]) {
    if (!state || !out || noutput <= 0) return io_fail(MCG_E_INVALID, "bad arguments");
    const unsigned sample_count = unsigned(step) / unsigned(noutput);
    const float count = float(sample_count);
    state[0] = (state[0] * count + tke_total) / (count + 1.0f);
    float abs_delta = 0.0f;
    if (state[2] != 0.0f) {
        state[2] = 0.0f;
    } else {
        abs_delta = std::fabs(state[0] - state[1]);
    }
    state[1] = state[0];
    out[0] = tke_total;
    out[1] = state[0];
    out[2] = abs_delta;
    return MCG_OK;
}

int mcg_tke_files_reset(const char *mean_dir) {
    if (!mean_dir) return io_fail(MCG_E_INVALID, "null argument");
    const char *nm[3] = {"tke_total.bin", "tke_avg.bin", "tke_diff.bin"};
    for (int k = 0; k < 3; ++k) {
        const std::string p = std::string(mean_dir) + "/" + nm[k];
        FILE *f = std::fopen(p.c_str(), "wb");
        if (!f) return io_fail(MCG_E_IO, "cannot create " + p);
        std::fclose(f);
    }
    return MCG_OK;
}

int mcg_append_tke(const char *mean_dir, const float vals[3]) {
    if (!mean_dir || !vals) return io_fail(MCG_E_INVALID, "null argument");
    const char *nm[3] = {"tke_total.bin", "tke_avg.bin", "tke_diff.bin"};
    for (int k = 0; k < 3; ++k) {
        const std::string p = std::string(mean_dir) + "/" + nm[k];
        FILE *f = std::fopen(p.c_str(), "ab");
        if (!f) return io_fail(MCG_E_IO, "cannot append " + p);
        std::fwrite(&vals[k], sizeof(float), 1, f);
        std::fclose(f);
    }
    return MCG_OK;
}

}  // extern "C"
