// JSON configuration for the C-ABI: the keys, defaults, value coercions and
// error messages of draugr_amg_config_from_json (src/cfunction_interface.jl:30-221,
// c_api/draugr_amg.h:58-89), plus this backend's own keys ("precision",
// "seed", "graphs").  A small recursive-descent parser; unknown keys are
// ignored as in the reference.
#include "types.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace draugr {

namespace {

struct JVal {
    enum Kind { Null, Bool, Num, Str, Obj, Arr } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;   // also the literal text of numbers
    std::map<std::string, std::shared_ptr<JVal>> obj;
    std::vector<std::shared_ptr<JVal>> arr;
};

struct Parser {
    const char* p;
    explicit Parser(const char* s) : p(s) {}
    [[noreturn]] void fail(const std::string& m) { throw Error("JSON parse error: " + m); }
    void ws() { while (*p && std::isspace((unsigned char)*p)) ++p; }
    std::shared_ptr<JVal> value() {
        ws();
        auto v = std::make_shared<JVal>();
        if (*p == '{') {
            v->kind = JVal::Obj;
            ++p;
            ws();
            if (*p == '}') { ++p; return v; }
            for (;;) {
                ws();
                if (*p != '"') fail("expected key string");
                std::string k = string();
                ws();
                if (*p != ':') fail("expected ':'");
                ++p;
                v->obj[k] = value();
                ws();
                if (*p == ',') { ++p; continue; }
                if (*p == '}') { ++p; break; }
                fail("expected ',' or '}'");
            }
        } else if (*p == '[') {
            v->kind = JVal::Arr;
            ++p;
            ws();
            if (*p == ']') { ++p; return v; }
            for (;;) {
                v->arr.push_back(value());
                ws();
                if (*p == ',') { ++p; continue; }
                if (*p == ']') { ++p; break; }
                fail("expected ',' or ']'");
            }
        } else if (*p == '"') {
            v->kind = JVal::Str;
            v->str = string();
        } else if (!std::strncmp(p, "true", 4)) {
            v->kind = JVal::Bool; v->b = true; p += 4; v->str = "true";
        } else if (!std::strncmp(p, "false", 5)) {
            v->kind = JVal::Bool; v->b = false; p += 5; v->str = "false";
        } else if (!std::strncmp(p, "null", 4)) {
            v->kind = JVal::Null; p += 4; v->str = "nothing";
        } else {
            const char* s = p;
            char* e = nullptr;
            v->num = std::strtod(p, &e);
            if (e == p) fail(std::string("unexpected character '") + *p + "'");
            p = e;
            v->kind = JVal::Num;
            v->str.assign(s, (size_t)(e - s));
        }
        return v;
    }
    std::string string() {
        std::string out;
        ++p;  // opening quote
        while (*p && *p != '"') {
            if (*p == '\\') {
                ++p;
                switch (*p) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': out += '?'; p += 4; break;
                    default: out += *p;
                }
                ++p;
            } else {
                out += *p++;
            }
        }
        if (*p != '"') fail("unterminated string");
        ++p;
        return out;
    }
};

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

double to_float(const JVal& v) {
    if (v.kind == JVal::Num) return v.num;
    if (v.kind == JVal::Bool) return v.b ? 1.0 : 0.0;
    char* e = nullptr;
    const double d = std::strtod(v.str.c_str(), &e);
    if (v.str.empty() || *e) throw Error("ArgumentError: cannot parse \"" + v.str + "\" as Float64");
    return d;
}
int to_int(const JVal& v) {
    if (v.kind == JVal::Num) return (int)std::lround(v.num);
    if (v.kind == JVal::Bool) return v.b ? 1 : 0;
    char* e = nullptr;
    const long d = std::strtol(v.str.c_str(), &e, 10);
    if (v.str.empty() || *e) throw Error("ArgumentError: invalid base 10 digit in \"" + v.str + "\"");
    return (int)d;
}
bool to_bool(const JVal& v) {
    if (v.kind == JVal::Bool) return v.b;
    const std::string s = lower(v.str);
    return s == "true" || s == "1";
}
std::string to_string(const JVal& v) { return lower(v.str); }

template <class T>
T lookup(const std::map<std::string, T>& table, const std::string& name, const char* label) {
    auto it = table.find(name);
    if (it != table.end()) return it->second;
    std::string valid;
    for (auto& kv : table) valid += (valid.empty() ? "" : ", ") + kv.first;
    throw Error(std::string("Unknown ") + label + ": \"" + name + "\". Valid: " + valid);
}

const std::map<std::string, Smoother> kSmoothers = {
    {"jacobi", Smoother::Jacobi}, {"colored_gs", Smoother::ColoredGS}, {"serial_gs", Smoother::SerialGS},
    {"spai0", Smoother::SPAI0}, {"spai1", Smoother::SPAI1}, {"l1_jacobi", Smoother::L1Jacobi},
    {"l1_colored_gs", Smoother::L1ColoredGS}, {"l1_serial_gs", Smoother::L1SerialGS},
    {"chebyshev", Smoother::Chebyshev}, {"ilu0", Smoother::ILU0}, {"serial_ilu0", Smoother::SerialILU0}};
const std::map<std::string, Coarsen> kCoarsen = {
    {"aggregation", Coarsen::Aggregation}, {"pmis", Coarsen::PMIS}, {"hmis", Coarsen::HMIS},
    {"rs", Coarsen::RS}, {"aggressive_pmis", Coarsen::AggressivePMIS},
    {"aggressive_hmis", Coarsen::AggressiveHMIS}, {"smoothed_aggregation", Coarsen::SmoothedAggregation}};
const std::map<std::string, Interp> kInterp = {
    {"direct", Interp::Direct}, {"standard", Interp::Standard}, {"extended_i", Interp::ExtendedI}};
const std::map<std::string, bool> kStrength = {{"absolute", false}, {"signed", true}};
const std::map<std::string, bool> kCycle = {{"v", false}, {"w", true}};
const std::map<std::string, Precision> kPrecision = {
    {"fp64", Precision::FP64}, {"fp32", Precision::FP32}, {"bf16", Precision::BF16}};

InterpSpec parse_interp(const JVal& v) {
    InterpSpec ip;  // defaults of the three constructors: trunc 0, max 0, norm_p 1, no rescale
    if (v.kind == JVal::Str) {
        ip.type = lookup(kInterp, lower(v.str), "interpolation");
        return ip;
    }
    if (v.kind != JVal::Obj) throw Error("\"interpolation\" must be a string or object");
    auto get = [&](const char* k) -> const JVal* {
        auto it = v.obj.find(k);
        return it == v.obj.end() ? nullptr : it->second.get();
    };
    const JVal* t = get("type");
    ip.type = lookup(kInterp, t ? to_string(*t) : std::string("extended_i"), "interpolation");
    if (const JVal* x = get("trunc_factor")) ip.trunc_factor = to_float(*x);
    if (ip.type == Interp::ExtendedI) {
        if (const JVal* x = get("max_elements")) ip.max_elements = to_int(*x);
        if (const JVal* x = get("norm_p")) ip.norm_p = to_int(*x);
        if (const JVal* x = get("rescale")) ip.rescale = to_bool(*x);
    }
    return ip;
}

CoarsenSpec make_coarsening(const std::string& name, double theta, const InterpSpec& ip) {
    CoarsenSpec c;
    c.type = lookup(kCoarsen, name, "coarsening");
    c.theta = theta;
    if (c.type == Coarsen::Aggregation || c.type == Coarsen::SmoothedAggregation) c.interp = InterpSpec();
    else c.interp = ip;
    return c;
}

}  // namespace

Config config_from_json(const char* json) {
    const std::string text = json ? json : "{}";
    Parser ps(text.c_str());
    auto root = ps.value();
    ps.ws();
    if (*ps.p) throw Error("JSON parse error: trailing characters");
    if (root->kind != JVal::Obj) throw Error("JSON parse error: top level must be an object");
    auto get = [&](const char* k) -> const JVal* {
        auto it = root->obj.find(k);
        return it == root->obj.end() ? nullptr : it->second.get();
    };
    Config cfg;  // AMGConfig() defaults
    InterpSpec ip;
    if (const JVal* v = get("interpolation")) ip = parse_interp(*v);
    double theta = 0.5;
    if (const JVal* v = get("theta")) theta = to_float(*v);
    cfg.coarsening = make_coarsening(get("coarsening") ? to_string(*get("coarsening")) : "hmis", theta, ip);
    // backend extension: the initial (first-levels) coarsening may carry its own
    // theta / interpolation ("initial_theta", "initial_interpolation")
    InterpSpec ip0 = ip;
    if (const JVal* v = get("initial_interpolation")) ip0 = parse_interp(*v);
    double theta0 = theta;
    if (const JVal* v = get("initial_theta")) theta0 = to_float(*v);
    cfg.initial_coarsening = get("initial_coarsening")
                                 ? make_coarsening(to_string(*get("initial_coarsening")), theta0, ip0)
                                 : cfg.coarsening;
    if (const JVal* v = get("sa_omega")) {
        cfg.coarsening.sa_omega = to_float(*v);
        cfg.initial_coarsening.sa_omega = cfg.coarsening.sa_omega;
    }
    if (const JVal* v = get("aggressive_base")) {
        // "aggregation": AggressiveCoarsening whose base is neither :pmis nor :hmis
        if (to_string(*v) == "aggregation") {
            if (cfg.coarsening.type == Coarsen::AggressivePMIS) cfg.coarsening.type = Coarsen::AggressiveAggregation;
            if (cfg.initial_coarsening.type == Coarsen::AggressivePMIS)
                cfg.initial_coarsening.type = Coarsen::AggressiveAggregation;
        }
    }
    bool omega_set = false;
    if (const JVal* v = get("smoother")) {
        if (v->kind == JVal::Str) {
            cfg.smoother = lookup(kSmoothers, lower(v->str), "smoother");
        } else if (v->kind == JVal::Obj) {
            auto it = v->obj.find("type");
            cfg.smoother = lookup(kSmoothers, it != v->obj.end() ? to_string(*it->second) : "l1_colored_gs",
                                  "smoother");
            auto om = v->obj.find("omega");
            if (om != v->obj.end()) { cfg.jacobi_omega = to_float(*om->second); omega_set = true; }
        } else {
            throw Error("\"smoother\" must be a string or object");
        }
    }
    if (!omega_set)
        if (const JVal* v = get("jacobi_omega")) cfg.jacobi_omega = to_float(*v);
    if (const JVal* v = get("strength")) cfg.signed_strength = lookup(kStrength, to_string(*v), "strength");
    if (const JVal* v = get("cycle")) cfg.w_cycle = lookup(kCycle, to_string(*v), "cycle");
    if (const JVal* v = get("max_levels")) cfg.max_levels = to_int(*v);
    if (const JVal* v = get("max_coarse_size")) cfg.max_coarse_size = to_int(*v);
    if (const JVal* v = get("pre_smoothing_steps")) cfg.pre_smoothing_steps = to_int(*v);
    if (const JVal* v = get("post_smoothing_steps")) cfg.post_smoothing_steps = to_int(*v);
    if (const JVal* v = get("verbose")) cfg.verbose = to_int(*v);
    if (const JVal* v = get("initial_coarsening_levels")) cfg.initial_coarsening_levels = to_int(*v);
    if (const JVal* v = get("max_row_sum")) cfg.max_row_sum = to_float(*v);
    if (const JVal* v = get("coarse_solve_on_cpu")) cfg.coarse_solve_on_cpu = to_bool(*v);
    // backend extensions
    if (const JVal* v = get("precision")) cfg.precision = lookup(kPrecision, to_string(*v), "precision");
    if (const JVal* v = get("seed")) cfg.seed = (uint64_t)std::llround(to_float(*v));
    if (const JVal* v = get("graphs")) {
        if (v->kind == JVal::Str) {
            static const std::map<std::string, int> kGraphs = {{"auto", -1}, {"true", 1}, {"false", 0},
                                                                {"1", 1},     {"0", 0}};
            cfg.graphs = lookup(kGraphs, to_string(*v), "graphs");
        } else {
            cfg.graphs = to_bool(*v) ? 1 : 0;
        }
    }
    if (const JVal* v = get("gs_symmetric")) cfg.gs_symmetric = to_bool(*v);
    if (const JVal* v = get("tail")) {
        static const std::map<std::string, int> kTail = {{"auto", -1}, {"off", 0}, {"on", 1}};
        cfg.tail = lookup(kTail, to_string(*v), "tail");
    }
    if (const JVal* v = get("batch_rhs_per_thread")) {
        cfg.batch_rpt = to_int(*v);
        if (!(cfg.batch_rpt == 0 || cfg.batch_rpt == 1 || cfg.batch_rpt == 4 || cfg.batch_rpt == 8 ||
              cfg.batch_rpt == 16 || cfg.batch_rpt == -1 || cfg.batch_rpt == -4))
            throw Error("batch_rhs_per_thread must be 0 (autotuned), 1, 4, 8, 16, -1 or -4 (LDS-staged 1 or 4)");
    }
    if (const JVal* v = get("gpu_setup")) cfg.gpu_setup = to_bool(*v);
    if (const JVal* v = get("fuse")) {
        static const std::map<std::string, int> kFuse = {{"auto", -1}, {"off", 0}, {"on", 1}};
        cfg.fuse = lookup(kFuse, to_string(*v), "fuse");
    }
    if (const JVal* v = get("gpu_setup_min_rows")) cfg.gpu_setup_min_rows = to_int(*v);
    if (const JVal* v = get("kernel")) {
        static const std::map<std::string, int> kKernel = {{"auto", -1}, {"stream", 0}, {"rowseq", 1}, {"pipe", 2},
                                                                 {"stream16", 3}, {"sorted", 4},
                                                                 {"stream16pipe", 5}, {"sortedpipe", 6},
                                                                 {"sortedquad", 7}, {"vector", 8},
                                                                 {"pattern", 9}};
        cfg.kernel = lookup(kKernel, to_string(*v), "kernel");
    }
    if (const JVal* v = get("filtering")) {
        cfg.coarsening.filtering = to_bool(*v);
        cfg.initial_coarsening.filtering = cfg.coarsening.filtering;
    }
    if (const JVal* v = get("filter_tol")) {
        cfg.coarsening.filter_tol = to_float(*v);
        cfg.initial_coarsening.filter_tol = cfg.coarsening.filter_tol;
    }
    if (cfg.max_levels < 1) throw Error("max_levels must be >= 1");
    if (cfg.pre_smoothing_steps < 0 || cfg.post_smoothing_steps < 0)
        throw Error("smoothing steps must be >= 0");
    return cfg;
}

}  // namespace draugr
