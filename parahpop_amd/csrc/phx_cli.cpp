// phx_gpu CONFIG.json TARGETS.json OUTPUT_DIRECTORY
//
// Process-level drop-in for the reference GPU executable (src/gpu/main.cpp:2 ->
// src/common/driver.cpp:57-186): same arguments, --help text, refusal of an
// output directory that already holds results, validation messages, output
// files (final_states.csv with 17 significant digits, summary.json,
// configuration_resolved.json) and exit status (0 success, 1 on any error with
// "ParaHPOP: <message>" on stderr).
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "json_lite.h"

using phx::json::Value;

namespace {

std::string absPath(const std::string& p) {
    if (!p.empty() && p[0] == '/') return p;
    char buf[PATH_MAX];
    if (!getcwd(buf, sizeof buf)) throw std::runtime_error("getcwd failed");
    return std::string(buf) + "/" + p;
}

std::string normalize(const std::string& p) {
    std::vector<std::string> parts;
    std::stringstream ss(p);
    std::string item;
    while (std::getline(ss, item, '/')) {
        if (item.empty() || item == ".") continue;
        if (item == "..") { if (!parts.empty()) parts.pop_back(); continue; }
        parts.push_back(item);
    }
    std::string out;
    for (auto& s : parts) out += "/" + s;
    return out.empty() ? "/" : out;
}

bool exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

std::string dirName(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? "." : (k == 0 ? "/" : p.substr(0, k));
}

// driver.cpp:41-55: rewrite documented path fields as absolute paths.
void resolvePaths(Value& v, const std::string& base, const std::string& key = "") {
    const bool pathField = key == "file" || key == "spaceWeatherFile" || key == "eopFile" || key == "constants" ||
                           key == "ephemeris" || key == "orientations";
    if (v.isString() && pathField) {
        if (v.str == "analytic") return;
        std::string p = v.str;
        if (p[0] != '/') p = base + "/" + p;
        v.str = normalize(p);
    } else if (v.isObject()) {
        for (auto& kv : *v.obj) resolvePaths(kv.second, base, kv.first);
    } else if (v.isArray()) {
        for (auto& item : *v.arr) resolvePaths(item, base, key);
    }
}

void mkdirs(const std::string& p) {
    std::string cur;
    std::stringstream ss(p);
    std::string item;
    while (std::getline(ss, item, '/')) {
        if (item.empty()) continue;
        cur += "/" + item;
        if (!exists(cur) && ::mkdir(cur.c_str(), 0755) != 0) throw std::runtime_error("Cannot create directory: " + cur);
    }
}

std::string g17(double x) {
    std::ostringstream o;
    o.precision(17);
    o << x;
    return o.str();
}

}  // namespace

int main(int argc, char** argv) {
    const char* backend = "device";
    phx_ctx* ctx = nullptr;
    try {
        if (argc == 2 && std::string(argv[1]) == "--help") {
            std::cout << "Usage: " << argv[0] << " CONFIG.json TARGETS.json OUTPUT_DIRECTORY\n"
                      << "Backend: " << backend << "; integrator: RKF78 fixed step\n";
            return 0;
        }
        if (argc != 4) throw std::runtime_error("Expected CONFIG.json TARGETS.json OUTPUT_DIRECTORY (or --help).");
        const std::string configPath = normalize(absPath(argv[1]));
        const std::string inputPath = normalize(absPath(argv[2]));
        const std::string outPath = normalize(absPath(argv[3]));
        for (const char* f : {"final_states.csv", "summary.json", "configuration_resolved.json"})
            if (exists(outPath + "/" + f)) throw std::runtime_error("Output exists; use a new directory: " + outPath);
        Value config = phx::json::readFile(configPath);
        if (!config.isObject()) throw std::runtime_error("configuration must be an object.");
        const double duration = config.at("run").at("duration_seconds").asDouble();
        if (!(duration > 0) || !std::isfinite(duration)) throw std::runtime_error("duration_seconds must be finite and positive.");
        const double step = config.at("execution").at("integration").at("initialStepSize").asDouble();
        // targets (driver.cpp:87-123)
        Value input = phx::json::readFile(inputPath);
        const Value& targets = input.at("targets");
        if (!targets.isArray() || targets.size() == 0) throw std::runtime_error("targets must be a nonempty array.");
        const size_t n = targets.size();
        std::vector<int64_t> id(n);
        std::vector<double> epoch(n), state(6 * n), mass(n), drag(n), srp(n), cd(n), cr(n);
        std::set<int64_t> ids;
        for (size_t i = 0; i < n; ++i) {
            const Value& t = targets.at(i);
            id[i] = t.at("id").asInt();
            if (!ids.insert(id[i]).second) throw std::runtime_error("Target IDs must be unique.");
            const Value& s = t.at("state_km_km_s");
            if (!s.isArray() || s.size() != 6) throw std::runtime_error("state_km_km_s must hold six numbers.");
            for (int d = 0; d < 6; ++d) {
                state[size_t(d) * n + i] = s.at(size_t(d)).asDouble();
                if (!std::isfinite(state[size_t(d) * n + i])) throw std::runtime_error("State contains a nonfinite value.");
            }
            epoch[i] = t.at("epoch_mjd2000_tdb").asDouble();
            mass[i] = t.at("mass_kg").asDouble();
            drag[i] = t.at("drag_area_m2").asDouble();
            srp[i] = t.at("srp_area_m2").asDouble();
            cd[i] = t.at("cd").asDouble();
            cr[i] = t.at("cr").asDouble();
        }
        const auto envStart = std::chrono::steady_clock::now();
        int rc = phx_create(configPath.c_str(), &ctx);
        if (rc != PHX_OK) throw std::runtime_error(phx_last_error(nullptr));
        const auto propStart = std::chrono::steady_clock::now();
        std::vector<double> outEpoch(n), outState(6 * n);
        phx_stats stats{};
        rc = phx_propagate(ctx, n, id.data(), epoch.data(), state.data(), mass.data(), drag.data(), srp.data(), cd.data(),
                           cr.data(), duration, outEpoch.data(), outState.data(), &stats);
        if (rc != PHX_OK) throw std::runtime_error(phx_last_error(ctx));
        const auto propEnd = std::chrono::steady_clock::now();
        // driver.cpp:145-155
        double maxErr = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double expected = epoch[i] + duration / 86400.0;
            const double err = std::fabs(outEpoch[i] - expected) * 86400.0;
            if (!std::isfinite(err) || err > 1e-4) throw std::runtime_error("A target did not reach its requested end epoch.");
            maxErr = std::max(maxErr, err);
            for (int d = 0; d < 6; ++d)
                if (!std::isfinite(outState[size_t(d) * n + i])) throw std::runtime_error("Nonfinite final state.");
        }
        mkdirs(outPath);
        {
            std::ofstream csv(outPath + "/final_states.csv");
            csv << "id,epoch_mjd2000_tdb,x_km,y_km,z_km,vx_km_s,vy_km_s,vz_km_s\n";
            for (size_t i = 0; i < n; ++i) {
                csv << id[i] << ',' << g17(outEpoch[i]);
                for (int d = 0; d < 6; ++d) csv << ',' << g17(outState[size_t(d) * n + i]);
                csv << '\n';
            }
            if (!csv) throw std::runtime_error("Failed to write final_states.csv.");
        }
        Value resolved = config;
        resolved.erase("run");
        Value ex = resolved.at("execution");
        ex.set("on", Value::of(std::string(backend)));
        resolved.set("execution", ex);
        resolvePaths(resolved, dirName(configPath));
        {
            std::ofstream r(outPath + "/configuration_resolved.json");
            r << phx::json::dump(phx::json::sortedKeys(resolved)) << '\n';
            if (!r) throw std::runtime_error("Failed to write resolved configuration.");
        }
        Value summary = Value::makeObject();
        summary.set("backend", Value::of(std::string(backend)));
        summary.set("targets", Value::ofInt((long long)n));
        summary.set("integrator", Value::of(std::string("RKF78 fixed step")));
        summary.set("initial_step_seconds", Value::of(step));
        summary.set("duration_seconds", Value::of(duration));
        summary.set("environment_seconds", Value::of(std::chrono::duration<double>(propStart - envStart).count()));
        summary.set("propagation_seconds", Value::of(std::chrono::duration<double>(propEnd - propStart).count()));
        summary.set("max_end_epoch_error_seconds", Value::of(maxErr));
        summary.set("validation_passed", Value::of(true));
        // the reference's nine keys, written as nlohmann::json orders them (sorted)
        summary = phx::json::sortedKeys(summary);
        {
            std::ofstream s(outPath + "/summary.json");
            s << phx::json::dump(summary) << '\n';
            if (!s) throw std::runtime_error("Failed to write summary.");
        }
        std::cout << phx::json::dump(summary) << '\n';
        phx_destroy(ctx);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ParaHPOP: " << e.what() << '\n';
        if (ctx) phx_destroy(ctx);
        return 1;
    }
}
