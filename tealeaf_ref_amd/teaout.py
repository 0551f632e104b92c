"""Reader for the ``tea.out`` log that the reference writes and the native host
reproduces line for line.

Line formats (reference): summary row ``'(a6,i7,5e26.17)'`` field_summary.f90:97;
step header timestep.f90:63; solver result lines tea_solve.f90:438-453; switch
and eigen lines tea_solve.f90:216-220; ``est itc`` tea_leaf_cheby.f90:245-250.
"""
import re

# numbers as Fortran writes them, including NaN and Infinity (a one-cell mesh
# has a zero residual and produces 0/0 in the reference as here)
_FLOAT = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[EeDd][-+]?\d+)?|[-+]?NaN|[-+]?Infinity|[-+]?Inf"
_STEP_HDR = re.compile(r"^\s*Step\s+(\d+)\s+time\s+(" + _FLOAT + r")\s+timestep\s+(" + _FLOAT + ")")
_SUMMARY = re.compile(r"^\s*step:\s*(\d+)\s+(" + _FLOAT + r")\s+(" + _FLOAT + r")\s+(" + _FLOAT +
                      r")\s+(" + _FLOAT + r")\s+(" + _FLOAT + r")")
_ITERS = re.compile(r"^\s*Iteration count\s+(\d+)")
_PPCG = re.compile(r"^\s*PPCG Iteration count\s+(-?\d+)\s+\(Total\s+(-?\d+)\)")
_CONDERR = re.compile(r"^\s*Conduction error\s+(" + _FLOAT + ")")
_EXACT = re.compile(r"^\s*EXACT error calculated as\s*(" + _FLOAT + ")")
_SWITCH = re.compile(r"^\s*Switching after\s+(\d+)\s+CG its, error\s+(" + _FLOAT + ")")
_EIGEN = re.compile(r"^\s*Eigen min\s*(" + _FLOAT + r")\s+Eigen max\s*(" + _FLOAT + r")")
_TIMEPERIT = re.compile(r"^\s*Solve Time\s+(" + _FLOAT + r")\s+Its\s+(\d+)\s+Time Per It\s+(" +
                        _FLOAT + ")")


def _f(s):
    return float(s.replace("D", "E").replace("d", "e"))


def parse(text):
    """Return ``{"summaries": [...], "steps": {step: {...}}}``.

    ``summaries`` holds every ``step:`` row in order as dicts with keys step,
    volume, mass, density, energy, temperature.  ``steps`` maps the timestep
    number to its solver lines (iterations, ppcg_inner, ppcg_total,
    conduction_error, exact_error, switch_n, eigmin, eigmax, est_itc,
    solve_time, time_per_it).
    """
    summaries, steps = [], {}
    cur = None
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i]
        m = _STEP_HDR.match(ln)
        if m:
            cur = int(m.group(1))
            steps[cur] = {"time": _f(m.group(2)), "dt": _f(m.group(3))}
            i += 1
            continue
        m = _SUMMARY.match(ln)
        if m:
            summaries.append({"step": int(m.group(1)), "volume": _f(m.group(2)),
                              "mass": _f(m.group(3)), "density": _f(m.group(4)),
                              "energy": _f(m.group(5)), "temperature": _f(m.group(6))})
            i += 1
            continue
        rec = steps.get(cur) if cur is not None else None
        if rec is not None:
            m = _ITERS.match(ln)
            if m:
                rec["iterations"] = int(m.group(1))
            m = _PPCG.match(ln)
            if m:
                rec["ppcg_inner"] = int(m.group(1))
                rec["ppcg_total"] = int(m.group(2))
            m = _CONDERR.match(ln)
            if m:
                rec["conduction_error"] = _f(m.group(1))
            m = _EXACT.match(ln)
            if m:
                rec["exact_error"] = _f(m.group(1))
            m = _SWITCH.match(ln)
            if m:
                rec["switch_n"] = int(m.group(1))
            m = _EIGEN.match(ln)
            if m:
                rec["eigmin"] = _f(m.group(1))
                rec["eigmax"] = _f(m.group(2))
            m = _TIMEPERIT.match(ln)
            if m:
                rec["solve_time"] = _f(m.group(1))
                rec["time_per_it"] = _f(m.group(3))
            if ln.strip() == "est itc" and i + 1 < len(lines):
                rec["est_itc"] = int(lines[i + 1].split()[0])
        i += 1
    return {"summaries": summaries, "steps": steps}


def last_summary_per_step(parsed):
    """Map step -> last ``step:`` row printed for that step."""
    out = {}
    for s in parsed["summaries"]:
        out[s["step"]] = s
    return out


_NUM_IN_LINE = re.compile(r"(?<![A-Za-z_\d.])(?:" + _FLOAT + r")(?![A-Za-z_])")


def skeleton(text):
    """The line layout of a tea.out: every number replaced by ``#`` and runs of
    blanks collapsed, so two logs of the same deck compare equal line by line
    whatever their values and timings."""
    out = []
    for line in text.splitlines():
        line = _NUM_IN_LINE.sub("#", line)
        out.append(" ".join(line.split()))
    return out
