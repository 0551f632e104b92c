"""Seeded tea.in deck generator for the benchmark and parity workloads.

Ambient state: density 100, energy 1e-4 on [0,10]^2.  Rectangles: lower-left
corner ~ U[0,10)^2, width/height ~ U[0.5,3], clipped to the domain; density
0.1, energy 25.  Circles: centre ~ U[0,10)^2, radius ~ U[0.3,1.5], density
0.1, energy ~ U[0.1,25].  numpy PCG64 with the given seed; coordinates are
rounded to 4 decimals and every line fits the reader's 100-character record
(parse.f90:154).  The seed is written into the deck as a comment.
"""
import numpy as np

MAX_LINE = 100


def _num(v, nd=4):
    s = f"{v:.{nd}f}".rstrip("0")
    if s.endswith("."):
        s += "0"
    return s


def states_for(seed, n_rect=3, n_circ=0, domain=10.0):
    """State lines (2..) for the seeded geometry."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_rect):
        x0, y0 = rng.uniform(0.0, domain, size=2)
        w, h = rng.uniform(0.5, 3.0, size=2)
        x1, y1 = min(x0 + w, domain), min(y0 + h, domain)
        out.append(("rectangle", 0.1, 25.0, x0, y0, x1, y1, None))
    for _ in range(n_circ):
        cx, cy = rng.uniform(0.0, domain, size=2)
        rad = rng.uniform(0.3, 1.5)
        e = rng.uniform(0.1, 25.0)
        out.append(("circle", 0.1, e, cx, cy, None, None, rad))
    return out


def make_deck(cells, solver="cg", seed=1, n_rect=3, n_circ=0, coefficient="density",
              eps=1e-15, max_iters=5000, end_step=10, dt=0.004, summary_frequency=1,
              precond=None, inner_steps=None, presteps=None, epslim=None, reflective=False,
              halo_depth=None, check_result=False, profiler=False, extra=(), y_cells=None):
    """Return the text of a tea.in deck."""
    nx = cells
    ny = cells if y_cells is None else y_cells
    lines = ["*tea", f"! seeded deck: numpy PCG64 seed {seed}, {n_rect} rectangles, {n_circ} circles",
             "state 1 density=100.0 energy=0.0001"]
    for i, (geo, rho, e, a, b, c, d, rad) in enumerate(states_for(seed, n_rect, n_circ), start=2):
        if geo == "rectangle":
            lines.append(f"state {i} density={_num(rho)} energy={_num(e)} geometry=rectangle "
                         f"xmin={_num(a)} ymin={_num(b)} xmax={_num(c)} ymax={_num(d)}")
        else:
            lines.append(f"state {i} density={_num(rho)} energy={_num(e)} geometry=circle "
                         f"xmin={_num(a)} ymin={_num(b)} radius={_num(rad)}")
    lines += [f"x_cells={nx}", f"y_cells={ny}", "xmin=0.0", "ymin=0.0", "xmax=10.0", "ymax=10.0",
              f"initial_timestep={_num(dt, 6)}", f"end_step={end_step}", f"tl_max_iters={max_iters}",
              f"tl_use_{solver}", f"tl_eps={eps:.1e}".replace("e-0", "e-"),
              f"summary_frequency={summary_frequency}"]
    lines.append("tl_coefficient_inverse_density" if coefficient == "inverse_density"
                 else "tl_coefficient_density")
    if precond:
        lines.append(f"tl_preconditioner_type={precond}")
    if inner_steps is not None:
        lines.append(f"tl_ppcg_inner_steps={inner_steps}")
    if presteps is not None:
        lines.append(f"tl_ch_cg_presteps={presteps}")
    if epslim is not None:
        lines.append(f"tl_ch_cg_epslim={epslim:.6e}")  # keeps the '.' the reader needs
    if reflective:
        lines.append("reflective_boundary")
    if halo_depth:
        lines.append(f"halo_depth={halo_depth}")
    if check_result:
        lines.append("tl_check_result")
    if profiler:
        lines.append("profiler_on")
    lines += list(extra)
    lines.append("*endtea")
    for ln in lines:
        if len(ln) > MAX_LINE:
            raise ValueError(f"deck line longer than {MAX_LINE} characters: {ln!r}")
    return "\n".join(lines) + "\n"


# The benchmark workloads (see bench.py): 4096^2 deck of rectangles and circles.
def bench_deck(cells=4096, solver="cg", seed=2, **kw):
    return make_deck(cells, solver=solver, seed=seed, n_rect=3, n_circ=2, **kw)
