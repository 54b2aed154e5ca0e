"""LF_DEM parameter files and unit resolution, for the Stokes-resistance path.

A parameter file is a list of `keyword = value;` statements with `/* */`
comments (SimulationInit.cpp:891-939).  Forces and times carry a unit suffix
(`kn = 5000r`, `kt = 0.5kn`, `contact_relaxation_time = 1e-3h`,
global.h:72-82, Simulation.cpp:491-510); LF_DEM resolves every force into the
simulation's internal unit from the control option (`-r 0.2r`: rate 0.2 in
repulsive units; `-s 5r`: stress 5 in repulsive units), SimulationInit.cpp:
156-236, 286-356, 358-375, 447-472.  This module restates that procedure and
returns the parameters of this library (`lfd_params`, in the internal unit
the reference would use: hydrodynamic for rate control, the stress unit for
stress control) together with the run-length and output settings.

Behaviour kept from the reference:
  * defaults first (SimulationInit.cpp:941-1014), then the file's statements;
  * statements end with ';' and everything after the last ';' is ignored;
    whitespace is removed; comment handling removes everything up to each
    '*/' while a '/*' remains; '//' is an error;
  * unknown keywords are errors (`legacy_keys=True` maps the renamed
    `repulsiveforce_length` of LF_DEM's own check files to `repulsive_length`);
  * a force or time without a suffix is an error; `h` times of time_end and
    the output intervals are strains (SimulationInit.cpp:498-521);
  * forces are resolved iteratively from the unit force, each one once its
    unit's ratio is known; an unresolvable unit is an error.

Parameters of features outside this library (Brownian motion, cohesion,
critical load, rolling friction, walls, auto kn/kt, 2D, extensional flow)
raise UnsupportedParameter when they are switched on.
"""
import math


class ParameterError(ValueError):
    """The file or a value cannot be interpreted (LF_DEM throws runtime_error)."""


class UnsupportedParameter(ParameterError):
    """A feature that the Stokes-resistance path of this library does not have."""


UNIT_LONGNAME = {"h": "hydro", "r": "repulsion", "b": "brownian", "c": "cohesion", "cl": "critical_load",
                 "ft": "ft_max", "kn": "kn", "kt": "kt", "kr": "kr", "s": "stress"}

FORCE = {"repulsion", "cohesion", "brownian", "critical_load", "kn", "kt", "kr", "ft_max", "min_kn_auto_det",
         "max_kn_auto_det", "min_kt_auto_det", "max_kt_auto_det"}
TIME = {"contact_relaxation_time", "contact_relaxation_time_tan", "time_end", "time_interval_output_config",
        "time_interval_output_data", "initial_log_time"}
STRAIN_TAGGED = ("time_interval_output_data", "time_interval_output_config", "time_end", "initial_log_time")
INT = {"simulation_mode", "friction_model", "integration_method", "nb_output_data_log_time",
       "nb_output_config_log_time", "np_fixed"}
BOOL = {"monolayer", "log_time_interval", "out_data_particle", "out_data_interaction", "out_data_vel_components",
        "origin_zero_flow", "auto_determine_knkt", "fixed_dt", "out_bond_order_parameter6", "out_binary_conf",
        "keep_input_strain"}
STR = {"lubrication_model", "event_handler", "out_particle_stress"}
FLOAT = {"repulsive_length", "repulsive_max_length", "vdW_coeffient", "vdW_singularity_cutoff",
         "lub_reduce_parameter", "disp_max", "dt_max", "dt_min", "lub_max_gap", "interaction_range", "sd_coeff",
         "dt", "Pe_switch", "mu_static", "mu_dynamic", "mu_rolling", "overlap_target", "disp_tan_target",
         "memory_strain_avg", "memory_strain_k", "start_adjust", "min_dt_auto_det", "max_dt_auto_det",
         "rest_threshold", "theta_shear", "strain_reversal", "brownian_relaxation_time"}
KNOWN = FORCE | TIME | INT | BOOL | STR | FLOAT
LEGACY = {"repulsiveforce_length": "repulsive_length"}


def _defaults(scale):
    """setDefaultParameters (SimulationInit.cpp:941-1014); `scale` is the
    suffix of the control option."""
    d = [("Pe_switch", "5"), ("dt", "1e-4"), ("disp_max", "1e-3"), ("dt_max", "-1"), ("monolayer", "false"),
         ("rest_threshold", "1e-4"), ("integration_method", "1"), ("np_fixed", "0"), ("sd_coeff", "1"),
         ("lubrication_model", "tangential"), ("friction_model", "1"), ("time_end", "10h"), ("lub_max_gap", "0.5"),
         ("interaction_range", "-1"), ("lub_reduce_parameter", "1e-3"),
         ("contact_relaxation_time", "1e-3" + scale), ("contact_relaxation_time_tan", "-1" + scale)]
    if UNIT_LONGNAME.get(scale, scale) != "kn":
        d += [("kn", "2000" + scale), ("min_kn_auto_det", "1000" + scale), ("max_kn_auto_det", "1000000" + scale)]
    if UNIT_LONGNAME.get(scale, scale) != "kt":
        d += [("kt", "0.5kn"), ("min_kt_auto_det", "1000" + scale), ("max_kt_auto_det", "1000000" + scale)]
    d += [("min_dt_auto_det", "1e-7"), ("max_dt_auto_det", "1e-3")]
    if UNIT_LONGNAME.get(scale, scale) != "kr":
        d += [("kr", "0kn")]
    d += [("auto_determine_knkt", "false"), ("overlap_target", "0.05"), ("disp_tan_target", "0.05"),
          ("memory_strain_avg", "0.01"), ("memory_strain_k", "0.02"), ("start_adjust", "0.2"),
          ("repulsive_length", "0.05"), ("repulsive_max_length", "-1"), ("vdW_coeffient", "-1"),
          ("vdW_singularity_cutoff", "0.1"), ("mu_static", "1"), ("mu_dynamic", "-1"), ("mu_rolling", "0"),
          ("time_interval_output_data", "1e-2h"), ("time_interval_output_config", "1e-1h"),
          ("log_time_interval", "false"), ("initial_log_time", "1e-4h"), ("nb_output_data_log_time", "100"),
          ("nb_output_config_log_time", "100"), ("origin_zero_flow", "true"), ("out_data_particle", "true"),
          ("out_data_interaction", "true"), ("out_particle_stress", ""), ("out_binary_conf", "false"),
          ("out_data_vel_components", "false"), ("fixed_dt", "false"), ("dt_min", "-1"), ("theta_shear", "0"),
          ("event_handler", ""), ("simulation_mode", "0"), ("keep_input_strain", "false"),
          ("brownian_relaxation_time", "1"), ("out_bond_order_parameter6", "false")]
    return d


def split_suffix(s):
    """getSuffix (global.h:72-82): the unit starts at the first lower-case
    letter other than 'e' (so '1e5h' is 1e5 in 'h')."""
    for k, ch in enumerate(s):
        if ch in "abcdfghijklmnopqrstuvwxyz":
            return s[:k], s[k:]
    return s, None


def _atof(s):
    """C atof: the longest numeric prefix, 0 when there is none."""
    best = 0.0
    for k in range(len(s), 0, -1):
        try:
            return float(s[:k])
        except ValueError:
            continue
    return best


def _atoi(s):
    digits = ""
    for k, ch in enumerate(s):
        if ch.isdigit() or (k == 0 and ch in "+-"):
            digits += ch
        else:
            break
    try:
        return int(digits)
    except ValueError:
        return 0


def _bool(key, s):
    if s == "true":
        return True
    if s == "false":
        return False
    raise ParameterError("%s: the value should be true or false, got %r" % (key, s))


def statements(text):
    """(keyword, value) statements of a parameter file, in order
    (readParameterFile, SimulationInit.cpp:891-939)."""
    parts = text.split(";")[:-1]          # what follows the last ';' is never read
    out = []
    for raw in parts:
        s = "".join(raw.split())
        while True:
            b = s.find("/*")
            if b < 0:
                break
            e = s.find("*/")
            if e < 0 or b > e:
                raise ParameterError("syntax error in the parameter file")
            s = s[e + 2:]
        if "//" in s:
            raise ParameterError("// is not the syntax to comment out. Use /* comment */")
        eq = s.find("=")
        if eq < 0:
            key, val = s, s
        else:
            key, val = s[:eq], s[eq + 1:]
        out.append((key, val))
    return out


class Parameters:
    """Raw keyword values: the defaults, then a file's statements."""

    def __init__(self, scale_suffix, legacy_keys=False):
        self.scale = scale_suffix
        self.legacy = legacy_keys
        self.values = {}
        for k, v in _defaults(scale_suffix):
            self.set(k, v)

    def set(self, key, value):
        if self.legacy and key in LEGACY:
            key = LEGACY[key]
        if key not in KNOWN:
            raise ParameterError("keyword %s is not associated with an parameter" % key)
        if key in FORCE or key in TIME:
            num, suf = split_suffix(value)
            if suf is None:
                raise ParameterError("no unit scale (suffix) provided for %s" % key)
            if suf not in UNIT_LONGNAME:
                raise ParameterError("input value \"%s\" has an unknown unit \"%s\"" % (key, suf))
            self.values[key] = (_atof(num), UNIT_LONGNAME[suf])
        elif key in INT:
            self.values[key] = _atoi(value)
        elif key in BOOL:
            self.values[key] = _bool(key, value)
        elif key in STR:
            self.values[key] = value.replace('"', "") if key != "lubrication_model" else value
        else:
            v = _atof(value)
            if key == "theta_shear":
                v *= math.pi / 180
            self.values[key] = v

    def read(self, path):
        with open(path) as f:
            for k, v in statements(f.read()):
                self.set(k, v)
        return self


class Setup:
    """Resolved parameters of one run."""

    def __init__(self):
        self.params = {}          # lfd_params fields (internal unit)
        self.control = "rate"
        self.rate = 1.0           # rate control: shear rate in the internal unit
        self.target_stress = 0.0  # stress control: internal stress unit (LF_DEM's input / (6 pi))
        self.internal_unit = "hydro"
        self.output_unit = "hydro"   # unit of the data files: the control value's (SimulationInit.cpp:471)
        self.output_ratio = 1.0      # internal force unit / output force unit (Simulation.cpp:654-660)
        self.strain_end = None    # time_end given as a strain
        self.time_end = None      # time_end given as a time (internal unit)
        self.output_data = None   # (value, "strain" | "time")
        self.output_config = None
        self.fixed_dt = False
        self.raw = None


def _resolve_forces(vals, unit_force, ratios):
    """resolveUnitSystem (SimulationInit.cpp:156-236)."""
    vals[unit_force] = (1.0, unit_force)
    for name, (v, unit) in vals.items():
        if name in FORCE or name == "hydro" or name == unit_force:
            ratios[name + "/" + unit] = v
    while True:
        unsolved = solved = False
        for name in list(vals):
            v, unit = vals[name]
            if unit in (unit_force, "strain"):
                continue
            key = unit + "/" + unit_force
            if key in ratios:
                if name in TIME:
                    vals[name] = (v / ratios[key], unit_force)
                else:
                    vals[name] = (v * ratios[key], unit_force)
                    ratios[name + "/" + unit_force] = vals[name][0]
                solved = True
            else:
                unsolved = True
        if not (unsolved and solved):
            break
    if unsolved:
        bad = [n for n, (v, u) in vals.items() if u not in (unit_force, "strain")]
        raise ParameterError("input values %s have units that cannot be resolved" % bad)


def resolve(params, control, value):
    """Resolved Setup from Parameters and the control option: control "rate"
    with value like "0.2r" (LF_DEM -r), or "stress" with value like "5r" (-s)."""
    num, suf = split_suffix(value)
    if suf is None or suf not in UNIT_LONGNAME:
        raise ParameterError("the control value needs a unit suffix: %r" % value)
    x = _atof(num)
    scale = UNIT_LONGNAME[suf]
    raw = dict(params.values)
    _check_supported(raw)
    vals = {k: v for k, v in raw.items() if k in FORCE or k in TIME}
    for name in STRAIN_TAGGED:                        # tagStrainParameters
        if vals[name][1] == "hydro":
            vals[name] = (vals[name][0], "strain")
    ratios = {}
    S = Setup()
    S.raw = raw
    if control == "rate":
        if scale in vals and scale != "hydro":
            raise ParameterError("redefinition of the rate (given both in the control option and in the "
                                 "parameter file with \"%s\" force)" % scale)
        if x == 0:
            raise UnsupportedParameter("zero shear rate (Brownian relaxation) is outside this library")
        vals[scale] = (1.0 / x, "hydro")              # setupNonDimensionalizationRateControlled
        ratios[scale + "/hydro"] = 1.0 / x
        _resolve_forces(vals, "hydro", ratios)
        S.internal_unit = "hydro"
        S.rate = vals["hydro"][0] if "hydro" in vals else 1.0
        S.control = "rate"
        S.output_unit = scale
        S.output_ratio = 1.0 / ratios[scale + "/hydro"] if scale != "hydro" else 1.0
    elif control == "stress":
        if scale == "hydro":
            raise ParameterError("please give a stress in non-hydro units")
        S.target_stress = x / 6 / math.pi
        for name, (v, u) in list(vals.items()):       # catchForcesInStressUnits
            if u == "stress":
                if name not in FORCE:
                    raise ParameterError("%s cannot be given in stress units" % name)
                vals[name] = (v * abs(S.target_stress), scale)
        _resolve_forces(vals, scale, ratios)
        S.internal_unit = scale
        S.rate = 1.0
        S.control = "stress"
        S.output_unit = scale
    else:
        raise ParameterError("unknown control variable %r" % control)
    te = vals["time_end"]
    if te[1] == "strain":
        S.strain_end = te[0]
    else:
        S.time_end = te[0]
    S.output_data = (vals["time_interval_output_data"][0],
                     "strain" if vals["time_interval_output_data"][1] == "strain" else "time")
    S.output_config = (vals["time_interval_output_config"][0],
                       "strain" if vals["time_interval_output_config"][1] == "strain" else "time")
    mu_s = raw["mu_static"]
    mu_d = raw["mu_dynamic"] if raw["mu_dynamic"] >= 0 else mu_s        # System.cpp:383-385
    lub = {"tangential": 2, "normal": 1}.get(raw["lubrication_model"])
    if lub is None:
        raise UnsupportedParameter("lubrication_model %r" % raw["lubrication_model"])
    rep = vals["repulsion"][0] if "repulsion" in vals else 0.0
    S.fixed_dt = raw["fixed_dt"]
    dt_min, dt_max = raw["dt_min"], raw["dt_max"]
    if S.fixed_dt:                                  # the constant step dt (System.cpp:607-610) is dt_max
        dt_min = dt_max = raw["dt"]
    S.params = dict(shear_rate=S.rate, theta_shear=raw["theta_shear"], lub_max_gap=raw["lub_max_gap"],
                    lub_reduce=raw["lub_reduce_parameter"], kn=vals["kn"][0], kt=vals["kt"][0],
                    mu_static=mu_s, mu_dynamic=mu_d, tau_n=vals["contact_relaxation_time"][0],
                    tau_t=vals["contact_relaxation_time_tan"][0], sd_coeff=raw["sd_coeff"], repulsion=rep,
                    repulsive_length=raw["repulsive_length"] if rep else 0.0, disp_max=raw["disp_max"],
                    dt_max=dt_max, dt_min=dt_min, friction_model=raw["friction_model"],
                    lubrication_model=lub, fixed_dt=int(bool(S.fixed_dt)))
    return S


def _check_supported(raw):
    if raw["integration_method"] != 1:
        raise UnsupportedParameter("integration_method %d (only the predictor-corrector, 1)"
                                   % raw["integration_method"])
    if raw["friction_model"] not in (0, 1):
        raise UnsupportedParameter("friction_model %d (0 frictionless, 1 Coulomb)" % raw["friction_model"])
    if raw["simulation_mode"] != 0:
        raise UnsupportedParameter("simulation_mode %d (walls and mixed problems)" % raw["simulation_mode"])
    if raw["np_fixed"] != 0:
        raise UnsupportedParameter("np_fixed (fixed particles)")
    if raw["monolayer"]:
        raise UnsupportedParameter("monolayer (2D)")
    if raw["auto_determine_knkt"]:
        raise UnsupportedParameter("auto_determine_knkt")
    if raw["mu_rolling"] > 0 or raw["kr"][0] != 0:
        raise UnsupportedParameter("rolling friction")
    for name in ("brownian", "cohesion", "critical_load", "ft_max"):
        if name in raw:
            raise UnsupportedParameter(name)
    if raw["vdW_coeffient"] > 0:
        raise UnsupportedParameter("van der Waals force")


def load(path, control, value, legacy_keys=False):
    """Setup of `LF_DEM -r value` (control="rate") or `-s value`
    (control="stress") with the parameter file `path`."""
    _, suf = split_suffix(value)
    return resolve(Parameters(suf or "h", legacy_keys=legacy_keys).read(path), control, value)
