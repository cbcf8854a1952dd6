// Frame output in the reference's wire formats (Source/fUtility.h):
//   .pos  write_fluid_particles (fUtility.h:254-303), write_solid_particles
//         (459-484), read_fluid_particles (486-510): four text lines, then
//         raw float32 records of the components named on line 2;
//   .json write_fluid_particles_json (390-427): positions and velocities
//         with y and z swapped, in jsoncpp's styled layout.
// Host code only; the arrays come from the library's snapshot.
#include "io.h"

#include "sphgpu.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace sphio {
namespace {

// component line, fUtility.h:69-98: "<" + " name " per component + ">"
std::string mask_string(unsigned comps) {
    std::string m = "<";
    if (comps & SPHGPU_POS_POSITION) m += " position ";
    if (comps & SPHGPU_POS_VELOCITY) m += " velocity ";
    if (comps & SPHGPU_POS_DENSITY) m += " density ";
    if (comps & SPHGPU_POS_PRESURE) m += " presure ";
    if (comps & SPHGPU_POS_ACCELERATION) m += " acceleration ";
    return m + ">";
}

// bytes per record named by a component line, fUtility.h:53-63
long long part_size(const std::string& mask) {
    long long s = 0;
    if (mask.find("position") != std::string::npos) s += 12;
    if (mask.find("velocity") != std::string::npos) s += 12;
    if (mask.find("color") != std::string::npos) s += 16;
    if (mask.find("density") != std::string::npos) s += 4;
    if (mask.find("presure") != std::string::npos) s += 4;
    if (mask.find("acceleration") != std::string::npos) s += 12;
    if (mask.find("dt") != std::string::npos) s += 4;
    return s;
}

// jsoncpp's valueToString(double): 17 significant digits, ".0" kept on
// integral values
std::string json_double(double v) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", v);
    std::string s(b);
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

}  // namespace

int write_pos(const char* path, const char* logName, long long frame, const char* kind, int index,
              long long n, unsigned comps, const float* pos, const float* vel, const float* rho,
              const float* prs, const float* acc) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::fprintf(f, "for parameters and scene info, please refer to \"%s\"\n", logName ? logName : "");
    std::fprintf(f, "components writed: %s\n", mask_string(comps).c_str());
    std::fprintf(f, "%lld (frame number), %s index: %d\n", frame, kind, index);
    std::fprintf(f, "%lld (num particles)\n", n);
    std::vector<float> rec;
    rec.reserve(16);
    for (long long i = 0; i < n; ++i) {
        rec.clear();
        if (comps & SPHGPU_POS_POSITION) rec.insert(rec.end(), pos + 3 * i, pos + 3 * i + 3);
        if (comps & SPHGPU_POS_VELOCITY) rec.insert(rec.end(), vel + 3 * i, vel + 3 * i + 3);
        if (comps & SPHGPU_POS_DENSITY) rec.push_back(rho[i]);
        if (comps & SPHGPU_POS_PRESURE) rec.push_back(prs[i]);
        if (comps & SPHGPU_POS_ACCELERATION) rec.insert(rec.end(), acc + 3 * i, acc + 3 * i + 3);
        if (std::fwrite(rec.data(), sizeof(float), rec.size(), f) != rec.size()) {
            std::fclose(f);
            return -2;
        }
    }
    return std::fclose(f) == 0 ? 0 : -2;
}

int write_json(const char* path, long long n, const float* pos, const float* vel, float d) {
    FILE* f = std::fopen(path, "w");
    if (!f) return -1;
    // members in jsoncpp's (sorted) order; with no particle only "num" exists
    std::fprintf(f, "{\n   \"num\" : %lld", n);
    if (n > 0) {
        auto triples = [&](const char* name, const float* a) {
            std::fprintf(f, ",\n   \"%s\" : [\n", name);
            for (long long i = 0; i < n; ++i) {
                // y and z swapped, fUtility.h:413-419
                std::fprintf(f, "      {\n         \"x\" : %s,\n         \"y\" : %s,\n         \"z\" : %s\n      }%s\n",
                             json_double(a[3 * i]).c_str(), json_double(a[3 * i + 2]).c_str(),
                             json_double(a[3 * i + 1]).c_str(), i + 1 < n ? "," : "");
            }
            std::fprintf(f, "   ]");
        };
        triples("position", pos);
        std::fprintf(f, ",\n   \"spacing_r\" : %s", json_double(d).c_str());
        std::fprintf(f, ",\n   \"temp_beta\" : [\n");
        for (long long i = 0; i < n; ++i)
            std::fprintf(f, "      {\n         \"value\" : 0.0\n      }%s\n", i + 1 < n ? "," : "");
        std::fprintf(f, "   ]");
        triples("velocity", vel);
    }
    std::fprintf(f, "\n}\n");
    return std::fclose(f) == 0 ? 0 : -2;
}

}  // namespace sphio

extern "C" int sphgpu_read_pos(const char* path, float* data, long long cap, long long* n, long long* frame,
                               char* mask, int mask_len) {
    if (!path || !n) return SPHGPU_E_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SPHGPU_E_ARG;
    char l1[512], l2[512], l3[512], l4[512];
    if (!std::fgets(l1, sizeof(l1), f) || !std::fgets(l2, sizeof(l2), f) || !std::fgets(l3, sizeof(l3), f) ||
        !std::fgets(l4, sizeof(l4), f)) {
        std::fclose(f);
        return SPHGPU_E_ARG;
    }
    long long num = -1, fr = -1;
    if (std::sscanf(l4, "%lld", &num) != 1 || num < 0) {
        std::fclose(f);
        return SPHGPU_E_ARG;
    }
    std::sscanf(l3, "%lld", &fr);
    std::string m(l2);
    while (!m.empty() && (m.back() == '\n' || m.back() == '\r')) m.pop_back();
    *n = num;
    if (frame) *frame = fr;
    if (mask && mask_len > 0) {
        std::strncpy(mask, m.c_str(), size_t(mask_len) - 1);
        mask[mask_len - 1] = '\0';
    }
    int rc = SPHGPU_OK;
    if (data) {
        const long long floats = num * (sphio::part_size(m) / 4);
        if (floats > cap) rc = SPHGPU_E_CAPACITY;
        else if (std::fread(data, sizeof(float), size_t(floats), f) != size_t(floats)) rc = SPHGPU_E_ARG;
    }
    std::fclose(f);
    return rc;
}
