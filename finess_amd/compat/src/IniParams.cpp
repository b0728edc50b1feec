// IniParams.cpp -- parameters.ini reader (see IniParams.h).
#include "IniParams.h"

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "../../../include/finess_hip.h"

IniParams global_ini_params;

namespace {
std::string rtrim(const std::string& s)
{
    size_t e = s.find_last_not_of(" \t\r");
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}
std::string ltrim(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t");
    return b == std::string::npos ? std::string() : s.substr(b);
}
bool is_name_char(char c, bool first, bool item)
{
    if (first) return std::isalpha((unsigned char)c) != 0;
    return std::isalnum((unsigned char)c) || c == '_' || (item && (c == '-' || c == '(' || c == ')'));
}
[[noreturn]] void terminate(const std::string& msg) { throw std::runtime_error(msg); }
} // namespace

void IniParams::init(const std::string& filename)
{
    std::ifstream f(filename.c_str());
    if (!f) terminate("cannot open " + filename);
    std::stringstream ss;
    ss << f.rdbuf();
    init_from_string(ss.str());
}

void IniParams::init_from_string(const std::string& text)
{
    doc.clear();
    std::istringstream in(text);
    std::string line, section;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::string s = ltrim(line);
        if (s.empty() || s[0] == ';') continue;
        if (s[0] == '[') {
            size_t e = s.find(']');
            if (e == std::string::npos) terminate("IniParser: bad section line " + std::to_string(lineno));
            std::string name = s.substr(1, e - 1);
            for (size_t k = 0; k < name.size(); k++)
                if (!is_name_char(name[k], k == 0, false)) terminate("IniParser: bad section name line " + std::to_string(lineno));
            std::string rest = ltrim(s.substr(e + 1));
            if (!rest.empty() && rest[0] != ';') terminate("IniParser: junk after section line " + std::to_string(lineno));
            section = name;
            continue;
        }
        size_t k = 0;
        while (k < s.size() && is_name_char(s[k], k == 0, true)) k++;
        if (k == 0) terminate("IniParser: syntax error line " + std::to_string(lineno));
        std::string key = s.substr(0, k);
        std::string rest = ltrim(s.substr(k));
        if (rest.empty() || rest[0] != '=') terminate("IniParser: expected '=' line " + std::to_string(lineno));
        std::string val = ltrim(rest.substr(1));
        size_t c = val.find(';');
        if (c != std::string::npos) val = val.substr(0, c);
        doc[section][key] = rtrim(val);  // last assignment wins
    }
    load();
}

std::string IniParams::raw(const char* sec, const char* key, const char* dflt, bool required) const
{
    auto s = doc.find(sec);
    if (s != doc.end()) {
        auto k = s->second.find(key);
        if (k != s->second.end()) return k->second;
    }
    if (required) terminate(std::string(sec) + "." + key + " is required");
    return dflt ? dflt : "";
}

const std::string& IniParams::get_gpu(const std::string& key, const std::string& dflt) const
{
    static thread_local std::string tmp;
    tmp = raw("gpu", key.c_str(), dflt.c_str(), false);
    return tmp;
}

namespace {
bool as_bool(const std::string& v, const std::string& name)
{
    if (v == "true" || v == "1") return true;
    if (v == "false" || v == "0") return false;
    terminate(name + " must be true or false");
}
int as_int(const std::string& v, const std::string& name)
{
    char* e = nullptr;
    long x = std::strtol(v.c_str(), &e, 10);
    if (e == v.c_str() || *ltrim(e).c_str()) terminate(name + " must be an integer");
    return (int)x;
}
double as_double(const std::string& v, const std::string& name)
{
    char* e = nullptr;
    double x = std::strtod(v.c_str(), &e);
    if (e == v.c_str() || *ltrim(e).c_str()) terminate(name + " must be a number");
    return x;
}
} // namespace

int IniParams::app_int(const char* sec, const char* key, const char* dflt) const
{
    return as_int(raw(sec, key, dflt, dflt == nullptr), std::string(sec) + "." + key);
}

double IniParams::app_double(const char* sec, const char* key, const char* dflt) const
{
    return as_double(raw(sec, key, dflt, dflt == nullptr), std::string(sec) + "." + key);
}

bool IniParams::app_bool(const char* sec, const char* key, const char* dflt) const
{
    return as_bool(raw(sec, key, dflt, dflt == nullptr), std::string(sec) + "." + key);
}

std::string IniParams::app_string(const char* sec, const char* key, const char* dflt) const
{
    return raw(sec, key, dflt, dflt == nullptr);
}

void IniParams::app_check_failed(const char* name, const char* rule) const
{
    terminate(std::string(name) + " should be " + rule);
}

void IniParams::load()
{
    // [finess]
    global_alpha = as_bool(raw("finess", "global_alpha", "false", false), "finess.global_alpha");
    mpp_limiter = as_bool(raw("finess", "mpp_limiter", "false", false), "finess.mpp_limiter");
    output_dir = raw("finess", "output_dir", "output", false);
    ndims = as_int(raw("finess", "ndims", nullptr, true), "finess.ndims");
    nout = as_int(raw("finess", "nout", "1", false), "finess.nout");
    tfinal = as_double(raw("finess", "tfinal", nullptr, true), "finess.tfinal");
    initial_dt = as_double(raw("finess", "initial_dt", nullptr, true), "finess.initial_dt");
    max_dt = as_double(raw("finess", "max_dt", nullptr, true), "finess.max_dt");
    desired_cfl = as_double(raw("finess", "desired_cfl", nullptr, true), "finess.desired_cfl");
    max_cfl = as_double(raw("finess", "max_cfl", nullptr, true), "finess.max_cfl");
    nv = as_int(raw("finess", "nv", nullptr, true), "finess.nv");
    const std::string tsm = raw("finess", "time_stepping_method", nullptr, true);
    if (tsm == "Runge-Kutta") time_stepping_method = TimeSteppingMethod::RK;
    else if (tsm == "Lax-Wendroff") time_stepping_method = TimeSteppingMethod::LxW;
    else if (tsm == "Multiderivative") time_stepping_method = TimeSteppingMethod::MD;
    else if (tsm == "SDC") time_stepping_method = TimeSteppingMethod::SDC;
    else if (tsm == "User-Defined") time_stepping_method = TimeSteppingMethod::USER_DEFINED;
    else terminate("finess.time_stepping_method: unknown value " + tsm);
    space_order = as_int(raw("finess", "space_order", nullptr, true), "finess.space_order");
    time_order = as_int(raw("finess", "time_order", nullptr, true), "finess.time_order");
    verbosity = as_int(raw("finess", "verbosity", nullptr, true), "finess.verbosity");
    mcapa = as_int(raw("finess", "mcapa", "0", false), "finess.mcapa");
    maux = as_int(raw("finess", "maux", nullptr, true), "finess.maux");
    source_term = as_bool(raw("finess", "source_term", nullptr, true), "finess.source_term");
    meqn = as_int(raw("finess", "meqn", nullptr, true), "finess.meqn");
    const std::string df = raw("finess", "datafmt", "ASCII", false);
    datafmt = df == "Silo" ? DataFormat::Silo : DataFormat::ASCII;
    // [weno]
    const std::string wv = raw("weno", "weno_version", "JS", false);
    if (wv == "JS") weno_version = WenoVersion::JS;
    else if (wv == "FD") weno_version = WenoVersion::FD;
    else if (wv == "Z") weno_version = WenoVersion::Z;
    else terminate("weno.weno_version: unknown value " + wv);
    power_param = as_double(raw("weno", "power_param", "2.0", false), "weno.power_param");
    alpha_scaling = as_double(raw("weno", "alpha_scaling", "1.1", false), "weno.alpha_scaling");
    epsilon = as_double(raw("weno", "epsilon", "1e-6", false), "weno.epsilon");
    // [grid]
    mx = as_int(raw("grid", "mx", nullptr, true), "grid.mx");
    mbc = as_int(raw("grid", "mbc", nullptr, true), "grid.mbc");
    xlow = as_double(raw("grid", "xlow", nullptr, true), "grid.xlow");
    xhigh = as_double(raw("grid", "xhigh", nullptr, true), "grid.xhigh");
    dx = (xhigh - xlow) / mx;
    if (ndims >= 2) {
        my = as_int(raw("grid", "my", nullptr, true), "grid.my");
        ylow = as_double(raw("grid", "ylow", nullptr, true), "grid.ylow");
        yhigh = as_double(raw("grid", "yhigh", nullptr, true), "grid.yhigh");
        dy = (yhigh - ylow) / my;
    }
    gamma = as_double(raw("euler", "gamma", "1.4", false), "euler.gamma");
    // checks (finess.py, weno.py)
    if (!(ndims == 1 || ndims == 2 || ndims == 3)) terminate("finess.ndims must be 1, 2 or 3");
    if (!(nout > 0)) terminate("finess.nout should > 0");
    if (!(tfinal >= 0)) terminate("finess.tfinal should >= 0");
    if (!(initial_dt > 0)) terminate("finess.initial_dt should > 0");
    if (!(max_dt >= initial_dt)) terminate("finess.max_dt should > finess.initial_dt");
    if (!(desired_cfl > 0)) terminate("finess.desired_cfl should > 0");
    if (!(max_cfl >= desired_cfl)) terminate("finess.max_cfl should > this->desired_cfl");
    if (!(nv > 0)) terminate("finess.nv should > 0");
    if (!(space_order == 1 || space_order == 3 || space_order == 5 || space_order == 7 || space_order == 9 ||
          space_order == 11))
        terminate("finess.space_order must be one of 1,3,5,7,9,11");
    if (!(time_order >= 1 && time_order <= 5)) terminate("finess.time_order must be one of 1..5");
    if (!(verbosity == 0 || verbosity == 1)) terminate("finess.verbosity must be 0 or 1");
    if (!(maux >= mcapa)) terminate("finess.maux should >= finess.mcapa");
    if (!(meqn >= 1)) terminate("finess.meqn should >= 1");
    if (!(alpha_scaling >= 1.0)) terminate("weno.alpha_scaling should >= 1");
    if (!(epsilon > 0)) terminate("weno.epsilon should > 0");
}

std::string IniParams::ini_doc_as_string() const
{
    std::ostringstream o;
    for (const auto& s : doc) {
        o << "[" << s.first << "]\n";
        for (const auto& kv : s.second) o << kv.first << " = " << kv.second << "\n";
    }
    return o.str();
}

void IniParams::to_finess_params(finess_params* p) const
{
    finess_params_default(p);
    p->ndims = ndims;
    p->mx = mx;
    p->my = ndims >= 2 ? my : 1;
    p->mbc = mbc;
    p->xlow = xlow; p->xhigh = xhigh; p->ylow = ylow; p->yhigh = yhigh;
    p->meqn = meqn;
    p->maux = maux;
    p->space_order = space_order;
    p->time_order = time_order;
    // User-Defined is the stepper the app links as FinSolveUser; the
    // reference's apps link FinSolveRK_PIF (time_order 4), as this library's
    // FinSolveUser does
    p->time_stepping_method = time_stepping_method == TimeSteppingMethod::LxW ? FINESS_TS_LXW
                              : (time_stepping_method == TimeSteppingMethod::USER_DEFINED && time_order == 4)
                                  ? FINESS_TS_PIF
                                  : FINESS_TS_RK;
    p->initial_dt = initial_dt;
    p->max_dt = max_dt;
    p->desired_cfl = desired_cfl;
    p->max_cfl = max_cfl;
    p->nv = nv;
    p->global_alpha = global_alpha;
    p->source_term = source_term;
    p->mpp_limiter = mpp_limiter ? 1 : 0;
    p->verbosity = verbosity;
    p->weno_version = weno_version == WenoVersion::FD ? FINESS_WENO_FD
                      : weno_version == WenoVersion::Z ? FINESS_WENO_Z : FINESS_WENO_JS;
    p->epsilon = epsilon;
    p->power_param = power_param;
    p->alpha_scaling = alpha_scaling;
    p->gamma = gamma;
    const std::string prec = raw("gpu", "precision", "fp64", false);
    p->precision = prec == "fp32" ? FINESS_PREC_FP32 : FINESS_PREC_FP64;
    const std::string rt = raw("gpu", "routing", "none", false);
    p->routing = rt == "learned" ? FINESS_ROUTE_LEARNED : rt == "oracle" ? FINESS_ROUTE_ORACLE : FINESS_ROUTE_NONE;
    p->route_tau = as_double(raw("gpu", "route_tau", "0.05", false), "gpu.route_tau");
    p->sw_compat_hl = as_bool(raw("gpu", "sw_compat_hl", "true", false), "gpu.sw_compat_hl");
    p->route_per_step = as_int(raw("gpu", "route_per_step", "1", false), "gpu.route_per_step");
    p->host_controller = as_bool(raw("gpu", "host_controller", "false", false), "gpu.host_controller") ? 1 : 0;
    // the device form of SourceTermFunc when the app's is not linked (the
    // drop-in probe identifies it from the app's own callback otherwise)
    const std::string sk = raw("gpu", "source_kind", "zero", false);
    if (sk == "zero") p->source_kind = FINESS_SRC_ZERO;
    else if (sk == "sw_bathymetry") p->source_kind = FINESS_SRC_SW_BATHYMETRY;
    else if (sk == "damping") p->source_kind = FINESS_SRC_DAMPING;
    else throw std::runtime_error("gpu.source_kind: unknown source term form " + sk);
    p->source_coef = as_double(raw("gpu", "source_coef", "0", false), "gpu.source_coef");
}
