// IniParams.h -- typed access to parameters.ini for FINESS apps, written for
// this library.  The accessor names, defaults, enum scopes and checks follow
// the reference's parameter schema (python/finess/params/finess.py:13-170,
// weno.py:15-51, dim1.py, dim2.py:8-65, euler.py:13-19) and the INI grammar
// of lib/sys/IniParser.rl:56-86 (sections [name], `key = value`, ';'
// comments, trailing spaces trimmed, the last assignment wins).  A [gpu]
// section adds the device-path options.
#ifndef FINESS_COMPAT_INIPARAMS_H
#define FINESS_COMPAT_INIPARAMS_H

#include <map>
#include <string>

struct finess_params;

class IniParams {
public:
    struct WenoVersion { enum enum_type { FD, JS, Z }; };
    struct TimeSteppingMethod { enum enum_type { RK, LxW, MD, SDC, USER_DEFINED }; };
    struct DataFormat { enum enum_type { ASCII, Silo }; };

    // Parses the file; throws std::runtime_error on a syntax error, a missing
    // required parameter or a failed check (the reference's terminate()).
    void init(const std::string& filename);
    void init_from_string(const std::string& text);
    std::string ini_doc_as_string() const;

    // [finess]
    bool get_global_alpha() const { return global_alpha; }
    bool get_mpp_limiter() const { return mpp_limiter; }
    const std::string& get_output_dir() const { return output_dir; }
    int get_ndims() const { return ndims; }
    int get_nout() const { return nout; }
    double get_tfinal() const { return tfinal; }
    double get_initial_dt() const { return initial_dt; }
    double get_max_dt() const { return max_dt; }
    double get_desired_cfl() const { return desired_cfl; }
    double get_max_cfl() const { return max_cfl; }
    int get_nv() const { return nv; }
    TimeSteppingMethod::enum_type get_time_stepping_method() const { return time_stepping_method; }
    int get_space_order() const { return space_order; }
    int get_time_order() const { return time_order; }
    int get_verbosity() const { return verbosity; }
    int get_mcapa() const { return mcapa; }
    int get_maux() const { return maux; }
    bool get_source_term() const { return source_term; }
    int get_meqn() const { return meqn; }
    DataFormat::enum_type get_datafmt() const { return datafmt; }
    // [weno]
    WenoVersion::enum_type get_weno_version() const { return weno_version; }
    double get_power_param() const { return power_param; }
    double get_alpha_scaling() const { return alpha_scaling; }
    double get_epsilon() const { return epsilon; }
    // [grid]
    int get_mx() const { return mx; }
    int get_my() const { return my; }
    int get_mbc() const { return mbc; }
    double get_xlow() const { return xlow; }
    double get_xhigh() const { return xhigh; }
    double get_ylow() const { return ylow; }
    double get_yhigh() const { return yhigh; }
    double get_dx() const { return dx; }
    double get_dy() const { return dy; }
    // [euler]
    double get_gamma() const { return gamma; }
    // [gpu] (no counterpart in the reference)
    const std::string& get_gpu(const std::string& key, const std::string& dflt) const;

    // Typed values of an app's own parameters.  The reference generates one
    // accessor per parameter an app declares in its this_app_params.py; here
    // gen_app_params.py (next to this library) writes those accessors into a
    // file that an app build names with -DFINESS_APP_PARAMS="<file>", and
    // they read their values through these: a missing key terminates (throws
    // std::runtime_error) unless the declaration gives a default.
    int app_int(const char* sec, const char* key, const char* dflt) const;
    double app_double(const char* sec, const char* key, const char* dflt) const;
    bool app_bool(const char* sec, const char* key, const char* dflt) const;
    std::string app_string(const char* sec, const char* key, const char* dflt) const;
    [[noreturn]] void app_check_failed(const char* name, const char* rule) const;
#ifdef FINESS_APP_PARAMS
#include FINESS_APP_PARAMS
#endif

    // The C-ABI parameter block (include/finess_hip.h) for these settings;
    // physics and boundary fields are left for the caller to fill.
    void to_finess_params(finess_params* p) const;

private:
    std::map<std::string, std::map<std::string, std::string> > doc;
    std::string raw(const char* sec, const char* key, const char* dflt, bool required) const;
    void load();

    bool global_alpha = false, mpp_limiter = false, source_term = false;
    std::string output_dir = "output";
    int ndims = 2, nout = 1, nv = 0, space_order = 5, time_order = 3, verbosity = 0, mcapa = 0, maux = 0, meqn = 1;
    double tfinal = 0, initial_dt = 0, max_dt = 0, desired_cfl = 0, max_cfl = 0;
    TimeSteppingMethod::enum_type time_stepping_method = TimeSteppingMethod::RK;
    DataFormat::enum_type datafmt = DataFormat::ASCII;
    WenoVersion::enum_type weno_version = WenoVersion::JS;
    double power_param = 2.0, alpha_scaling = 1.1, epsilon = 1e-6;
    int mx = 1, my = 1, mbc = 3;
    double xlow = 0, xhigh = 1, ylow = 0, yhigh = 1, dx = 1, dy = 1;
    double gamma = 1.4;
};

extern IniParams global_ini_params;

#endif
