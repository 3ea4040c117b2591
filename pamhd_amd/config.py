"""The reference's JSON run configuration (tests/mhd/config_files/**.json,
read by tests/mhd/test.cpp:180-290) for the grids this library steps:
periodic in every dimension longer than one cell, no boundaries, no
adaptive refinement.

    cfg = config.load("blast_wave.json")
    state, bg = cfg.initial_state()            # (8, n) face-once state, (9, n) or None
    result = config.run(cfg)                   # test.cpp's program loop on the GPU

Keys (defaults as the reference's):
* source/simulation_options.hpp:46-290: time-start (0), time-length (1),
  adiabatic-index (5/3), vacuum-permeability (4e-7 pi), proton-mass
  (1.672621777e-27), substep-min (0) / substep-max (999) as a non-negative
  integer or an expression of t, x, y, z, radius, lat, lon;
* source/mhd/options.hpp:40-117: save-mhd-n, solver-mhd,
  mhd-time-step-factor (0.5);
* source/grid/options.hpp:100-290: grid-options periodic / cells / volume /
  start expressions (`cells` is a variable of volume and start, `volume`
  of start);
* source/boundaries/geometries.hpp:106-137 with box.hpp / sphere.hpp:
  geometries referenced by initial conditions;
* source/boundaries/initial_conditions.hpp:84-125 and common.hpp:340-378:
  number-density, velocity, pressure, magnetic-field, each a "default"
  (number, array or expression) and optional "initial-conditions"
  regions {"geometry-id", "value"};
* source/background_magnetic_field.hpp:90-325: background-magnetic-field
  "value", "min-distance", "line-dipoles" and "dipoles";
* value-boundaries (time series) and copy-boundaries of each variable
  (source/boundaries/value_boundaries.hpp, copy_boundaries.hpp), with the
  grid's closed dimensions (pamhd_amd.boundary).
Anything else that changes the step (AMR, point-cloud data, two fluid
populations) raises ConfigError.

The initial state follows test.cpp:420-450: initialize_magnetic_field
(source/mhd/initialize.hpp:316-597: 7-point Gauss-Legendre averages of the
field over every face copy of every cell, with the emulated geometry of
source/grid/amr.hpp:288-310 for one-cell dimensions), update_vol_B from
each cell's own face copies, initialize_fluid (initialize.hpp:67-289), the
energy refit of apply_boundaries (source/mhd/boundaries.hpp:1261-1297 ->
update_vol_B with constant pressure, solve.hpp:1625-1672) and
sync_magnetic_field (source/common_functions.hpp:184-271), then keeps each
face once (the minus-side copy of the cell above it).
"""
import json
import math
import os

import numpy as np

from .expr import Expression, ExpressionError, _libm1

__all__ = ["ConfigError", "Config", "load", "run"]


class ConfigError(ValueError):
    pass


# Gauss-Legendre 7-point nodes and weights on [-1, 1] (textbook values)
_GL_X = (0.0, 0.405845151377397166906606412076961,
         0.741531185599394439863864773280788, 0.949107912342758524526189684047851)
_GL_W = (0.417959183673469387755102040816327, 0.381830050505118944950369775488975,
         0.279705391489276667901467771423780, 0.129484966168869693270611432679082)

_SOLVERS = ("rusanov", "hll-athena", "hlld-athena", "roe-athena")
_asin = _libm1(math.asin, np.arcsin)
_atan2 = np.frompyfunc(math.atan2, 2, 1)


def _gl7(f, a, b):
    """boost::math::quadrature::gauss<double, 7>::integrate(f, a, b):
    scale * (f(avg) w0 + sum_i (f(avg + scale x_i) + f(avg - scale x_i)) w_i)
    with avg = (a + b) / 2, scale = (b - a) / 2, element-wise over arrays"""
    avg = (a + b) * 0.5
    scale = (b - a) * 0.5
    r = f(avg + scale * 0.0) * _GL_W[0]
    for x, w in zip(_GL_X[1:], _GL_W[1:]):
        fp = f(avg + scale * x)
        fm = f(avg + scale * (-x))
        r = r + (fp + fm) * w
    return scale * r


class _Value:
    """A number, a fixed vector or an expression (fill_variable_value_from_json,
    source/boundaries/common.hpp:340-378)."""

    def __init__(self, obj, ncomp, what):
        self.ncomp = ncomp
        self.expr = None
        self.const = None
        if isinstance(obj, bool):
            raise ConfigError("%s: invalid value %r" % (what, obj))
        if isinstance(obj, (int, float)):
            if ncomp != 1:
                raise ConfigError("%s: expected an array of %d numbers" % (what, ncomp))
            self.const = float(obj)
        elif isinstance(obj, list):
            if ncomp == 1 or len(obj) != ncomp or not all(
                    isinstance(v, (int, float)) and not isinstance(v, bool) for v in obj):
                raise ConfigError("%s: expected %s" % (
                    what, "a number" if ncomp == 1 else "an array of %d numbers" % ncomp))
            self.const = tuple(float(v) for v in obj)
        elif isinstance(obj, str):
            try:
                self.expr = Expression(obj)
            except ExpressionError as e:
                raise ConfigError("%s: %s" % (what, e)) from None
        elif isinstance(obj, dict):
            raise ConfigError("%s: point-cloud data is not supported" % what)
        else:
            raise ConfigError("%s: invalid value %r" % (what, obj))

    def needs(self):
        return set() if self.expr is None else self.expr.variables()

    def __call__(self, shape, **env):
        """(ncomp, *shape) float array (a scalar variable gives (1, *shape))"""
        if self.expr is None:
            v = self.const if self.ncomp > 1 else (self.const,)
        else:
            v = self.expr.evaluate(**env)
            if self.ncomp == 1:
                if isinstance(v, tuple):
                    raise ConfigError("%r: expected a scalar" % self.expr.text)
                v = (v,)
            elif not isinstance(v, tuple) or len(v) != self.ncomp:
                raise ConfigError("%r: expected a vector of %d" % (self.expr.text, self.ncomp))
        return np.stack([np.broadcast_to(np.asarray(c, dtype=np.float64), shape) for c in v])


class _Variable:
    """initial_conditions.hpp: a default and the regions of one variable"""

    def __init__(self, obj, ncomp, name):
        if not isinstance(obj, dict) or "default" not in obj:
            raise ConfigError("%s: object with a default key required" % name)
        self.default = _Value(obj["default"], ncomp, name + " default")
        # value_boundaries.hpp / value_boundary.hpp: time series per geometry
        self.value_boundaries = []
        for i, b in enumerate(obj.get("value-boundaries", [])):
            what = "%s value-boundaries[%d]" % (name, i)
            gid = b.get("geometry-id") if isinstance(b, dict) else None
            if not isinstance(gid, int) or isinstance(gid, bool) or gid < 0:
                raise ConfigError(what + ": geometry-id required")
            stamps = b.get("time-stamps")
            vals = b.get("values")
            if not isinstance(stamps, list) or not stamps:
                raise ConfigError(what + ": time-stamps array required")
            if isinstance(vals, dict):
                raise ConfigError(what + ": point-cloud data is not supported")
            if not isinstance(vals, list) or len(vals) != len(stamps):
                raise ConfigError(what + ": one value per time stamp required")
            stamps = [float(t) for t in stamps]
            if stamps != sorted(stamps):
                raise ConfigError(what + ": time stamps must not decrease")
            self.value_boundaries.append((gid, stamps, [_Value(v, ncomp, what) for v in vals]))
        # copy_boundaries.hpp: geometry ids
        self.copy_boundaries = []
        for i, b in enumerate(obj.get("copy-boundaries", [])):
            gid = b.get("geometry-id") if isinstance(b, dict) else None
            if not isinstance(gid, int) or isinstance(gid, bool) or gid < 0:
                raise ConfigError("%s copy-boundaries[%d]: geometry-id required" % (name, i))
            self.copy_boundaries.append(gid)
        self.regions = []
        for i, r in enumerate(obj.get("initial-conditions", [])):
            gid = r.get("geometry-id") if isinstance(r, dict) else None
            if not isinstance(gid, int) or isinstance(gid, bool) or gid < 0 or "value" not in r:
                raise ConfigError("%s initial-conditions[%d]: geometry-id and value required" % (name, i))
            self.regions.append((gid, _Value(r["value"], ncomp, "%s region %d" % (name, i))))


def _number(obj, key, default, check=None):
    if key not in obj:
        return default
    v = obj[key]
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise ConfigError("JSON item %s is not a number" % key)
    v = float(v)
    if check is not None and not check(v):
        raise ConfigError("invalid %s: %r" % (key, v))
    return v


def _substep_option(obj, key, default):
    if key not in obj:
        return default
    v = obj[key]
    if isinstance(v, int) and not isinstance(v, bool):
        if v < 0:
            raise ConfigError("JSON item %s cannot be negative" % key)
        return v
    if isinstance(v, str):
        try:
            return Expression(v)
        except ExpressionError as e:
            raise ConfigError("%s: %s" % (key, e)) from None
    raise ConfigError("JSON item %s must either non-negative integer or string" % key)


class Config:
    def __init__(self, obj, path=None):
        if not isinstance(obj, dict):
            raise ConfigError("configuration must be a JSON object")
        self.path = path
        self.raw = obj
        second = [k for k in ("number-density2", "velocity2", "pressure2") if k in obj]
        if second:
            raise ConfigError("two-population configurations (%s) are not runs of the "
                              "single-fluid program" % ", ".join(second))
        normal = lambda v: math.isfinite(v) and v != 0 and abs(v) >= 2.2250738585072014e-308
        self.time_start = _number(obj, "time-start", 0.0, math.isfinite)
        self.time_length = _number(obj, "time-length", 1.0, lambda v: normal(v) and v > 0)
        self.adiabatic_index = _number(obj, "adiabatic-index", 5.0 / 3.0, lambda v: normal(v) and v > 0)
        self.vacuum_permeability = _number(obj, "vacuum-permeability", 4e-7 * math.pi,
                                           lambda v: normal(v) and v > 0)
        self.proton_mass = _number(obj, "proton-mass", 1.672621777e-27, lambda v: normal(v) and v > 0)
        self.substep_min = _substep_option(obj, "substep-min", 0)
        self.substep_max = _substep_option(obj, "substep-max", 999)
        self.output_directory = obj.get("output-directory", "")
        # source/mhd/options.hpp
        if "save-mhd-n" not in obj:
            raise ConfigError("JSON data doesn't have a save-mhd-n key")
        self.save_n = _number(obj, "save-mhd-n", -1.0)
        if "solver-mhd" not in obj:
            raise ConfigError("JSON data doesn't have a solver-mhd key")
        self.solver = obj["solver-mhd"]
        if self.solver not in _SOLVERS:
            raise ConfigError("unsupported solver: %s (tests/mhd/test.cpp runs %s)"
                              % (self.solver, ", ".join(_SOLVERS)))
        if "mhd-time-step-factor" not in obj:
            raise ConfigError("JSON data doesn't have a mhd-time-step-factor key")
        self.time_step_factor = _number(obj, "mhd-time-step-factor", 0.5)
        self._grid(obj)
        self._geometries(obj)
        self.number_density = _Variable(obj.get("number-density"), 1, "number-density")
        self.velocity = _Variable(obj.get("velocity"), 3, "velocity")
        self.pressure = _Variable(obj.get("pressure"), 1, "pressure")
        self.magnetic_field = _Variable(obj.get("magnetic-field"), 3, "magnetic-field")
        self._background(obj)

    def _background(self, obj):
        """background_magnetic_field.hpp:90-325: a constant, line dipoles
        and point dipoles, with min-distance (default: machine epsilon)"""
        self.background = None
        self.bg_min_distance = 2.220446049250313e-16
        self.bg_line_dipoles = []
        self.bg_dipoles = []
        if "background-magnetic-field" not in obj:
            return
        bg = obj["background-magnetic-field"]
        const = [0.0, 0.0, 0.0]
        if "value" in bg:
            v = bg["value"]
            if not isinstance(v, list) or len(v) != 3:
                raise ConfigError("Background magnetic field value must have 3 components")
            const = [float(c) for c in v]
        if "min-distance" in bg:
            self.bg_min_distance = float(bg["min-distance"])
        for i, ld in enumerate(bg.get("line-dipoles", [])):
            try:
                dim = int(ld["line-dimension"])
                fdir = int(ld["field-direction-through-line"])
                mag = float(ld["magnitude"])
                pos = [float(x) for x in ld["position"]]
            except (KeyError, TypeError, ValueError):
                raise ConfigError("line-dipoles[%d]: line-dimension, field-direction-through-line, "
                                  "magnitude and position required" % i) from None
            if not 1 <= dim <= 3 or fdir == 0 or abs(fdir) > 3 or len(pos) != 3:
                raise ConfigError("line-dipoles[%d]: invalid dimension, direction or position" % i)
            # the two coordinates across the line
            p2 = [pos[a] for a in range(3) if a != dim - 1]
            self.bg_line_dipoles.append((dim, fdir, mag, p2))
        for i, dp in enumerate(bg.get("dipoles", [])):
            try:
                m = [float(x) for x in dp["moment"]]
                pos = [float(x) for x in dp["position"]]
            except (KeyError, TypeError, ValueError):
                raise ConfigError("dipoles[%d]: moment and position required" % i) from None
            if len(m) != 3 or len(pos) != 3:
                raise ConfigError("dipoles[%d]: moment and position need 3 components" % i)
            self.bg_dipoles.append((m, pos))
        self.background = tuple(const)

    def background_field(self, x, y, z):
        """get_background_field (background_magnetic_field.hpp:331-494) at
        points (x, y, z): a (3, ...) array"""
        mu0 = self.vacuum_permeability
        shape = np.broadcast(x, y, z).shape
        P = [np.broadcast_to(np.asarray(a, dtype=np.float64), shape) for a in (x, y, z)]
        ret = [np.full(shape, c) for c in self.background]
        mind = self.bg_min_distance
        with np.errstate(all="ignore"):
            for dim, fdir, mag, p2 in self.bg_line_dipoles:
                dp = [None, None, None]
                others = [a for a in range(3) if a != dim - 1]
                dp[dim - 1] = P[dim - 1]
                dp[others[0]] = p2[0]
                dp[others[1]] = p2[1]
                r = [P[a] + (-dp[a]) for a in range(3)]
                r1 = np.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
                near = r1 < mind
                fnear = mu0 * mag / (2 * math.pi * math.pow(mind, 3))
                base = mag * mu0 / (2 * math.pi * r1 * r1 * r1 * r1)
                val = [base, base, base]
                ad = abs(fdir)
                if ad == dim:
                    for a in range(3):
                        val[a] = val[a] * (-r1 / 2) if a == dim - 1 else np.zeros(shape)
                    far = val
                else:
                    c1, c2 = others
                    v1, v2 = (c2, c1) if ad - 1 == c1 else (c1, c2)
                    val[dim - 1] = np.zeros(shape)
                    val[v1] = val[v1] * (r[c1] * r[c2])
                    val[v2] = val[v2] * (r[c2] * r[c2] - r[c1] * r[c1])
                    far = val
                if fdir < 0:
                    far = [v * -1 for v in far]
                # inside min-distance: the line-dimension (or the direction
                # through the line) component gets the clamped magnitude
                for a in range(3):
                    add_far = far[a]
                    if ad == dim:
                        add_near = fnear if a == dim - 1 else None
                        sign = 1
                    else:
                        add_near = fnear if a == ad - 1 else None
                        sign = 1 if fdir > 0 else -1
                    if add_near is None:
                        ret[a] = np.where(near, ret[a], ret[a] + add_far)
                    else:
                        nearv = ret[a] + add_near if sign > 0 else ret[a] - add_near
                        ret[a] = np.where(near, nearv, ret[a] + add_far)
            for m, pos in self.bg_dipoles:
                r = [P[a] + (-pos[a]) for a in range(3)]
                r1 = np.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
                near = r1 < mind
                fnear = mu0 / (2 * math.pi * math.pow(mind, 3))
                inv = 1 / r1
                u = [r[a] * inv for a in range(3)]
                dot = m[0] * u[0] + m[1] * u[1] + m[2] * u[2]
                f = mu0 / (4 * math.pi) / (r1 * r1 * r1)
                for a in range(3):
                    far = (u[a] * dot * 3 + (-m[a])) * f
                    ret[a] = np.where(near, ret[a] + m[a] * fnear, ret[a] + far)
        return np.stack(ret)

    # --- grid-options ------------------------------------------------------
    def _grid(self, obj):
        g = obj.get("grid-options")
        if not isinstance(g, dict):
            raise ConfigError("Object doesn't have a grid-options key")

        def vec3(key, env):
            if key not in g:
                raise ConfigError("grid-options doesn't have a %s key" % key)
            if not isinstance(g[key], str):
                raise ConfigError("Expression for grid %s isn't a string" % key)
            try:
                v = Expression(g[key]).evaluate(**env)
            except ExpressionError as e:
                raise ConfigError("grid-options %s: %s" % (key, e)) from None
            if not isinstance(v, tuple) or len(v) != 3:
                raise ConfigError("grid-options %s must be a vector of 3" % key)
            return v

        per = vec3("periodic", {})
        self.periodic = tuple(bool(np.asarray(p)) for p in per)
        # cells: GetFloat converted to the integral cell count
        cells = vec3("cells", {})
        self.cells = tuple(int(float(np.asarray(c))) for c in cells)
        if min(self.cells) < 1:
            raise ConfigError("grid-options cells must be >= 1: %r" % (self.cells,))
        cenv = {"cells": tuple(float(c) for c in self.cells)}
        self.volume = tuple(float(np.asarray(v)) for v in vec3("volume", cenv))
        cenv["volume"] = self.volume
        self.start = tuple(float(np.asarray(v)) for v in vec3("start", cenv))
        if g.get("max-ref-lvl", 0) not in (0,):
            raise ConfigError("adaptive mesh refinement (max-ref-lvl > 0) is not supported")
        # cell length, tests/mhd/test.cpp:339-346
        self.cell_length = tuple(self.volume[d] / self.cells[d] for d in range(3))

    def _geometries(self, obj):
        self.geometries = {}
        for i, gobj in enumerate(obj.get("geometries", [])):
            if "box" in gobj:
                b = gobj["box"]
                self.geometries[i] = ("box", tuple(map(float, b["start"])), tuple(map(float, b["end"])))
            elif "sphere" in gobj:
                s = gobj["sphere"]
                r = float(s["radius"])
                if r <= 0:
                    raise ConfigError("geometry %d: non-positive radius" % i)
                self.geometries[i] = ("sphere", tuple(map(float, s["center"])), r)
            else:
                raise ConfigError("geometries[%d] is not of a supported type" % i)

    # --- geometry of the cells ----------------------------------------------
    @property
    def n(self):
        return self.cells[0] * self.cells[1] * self.cells[2]

    def cell_bounds(self):
        """(min, max, centre) of every cell, (3, n) each, cell index
        i + nx (j + ny k): start + index x length, + length, + length / 2"""
        nx, ny, nz = self.cells
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        idx = (i.reshape(-1), j.reshape(-1), k.reshape(-1))
        lo = np.stack([self.start[d] + idx[d] * self.cell_length[d] for d in range(3)])
        hi = np.stack([lo[d] + self.cell_length[d] for d in range(3)])
        ctr = np.stack([lo[d] + self.cell_length[d] / 2 for d in range(3)])
        return lo, hi, ctr

    def emulated_bounds(self):
        """get_cell_geom_emulated: one-cell dimensions collapse to the grid's
        middle"""
        lo, hi, ctr = self.cell_bounds()
        lo, hi, ctr = lo.copy(), hi.copy(), ctr.copy()
        for d in range(3):
            if self.cells[d] == 1:
                end = self.start[d] + self.cell_length[d] * self.cells[d]
                mid = (end + self.start[d]) / 2
                lo[d] = hi[d] = ctr[d] = mid
        return lo, hi, ctr

    def overlapping(self, gid):
        """cells that overlap geometry gid (test.cpp:394-399 with the true
        cell bounds; box.hpp:128-155, sphere.hpp:125-153)"""
        if gid not in self.geometries:
            raise ConfigError("geometry-id %d does not exist" % gid)
        lo, hi, _ = self.cell_bounds()
        kind = self.geometries[gid]
        if kind[0] == "box":
            s, e = kind[1], kind[2]
            ok = np.ones(self.n, dtype=bool)
            for d in range(3):
                ok &= ~((lo[d] >= e[d]) | (hi[d] <= s[d]))
            return np.nonzero(ok)[0]
        c, r = kind[1], kind[2]
        dist2 = np.zeros(self.n)
        for d in range(3):
            below = c[d] < lo[d]
            above = c[d] > hi[d]
            dist2 = dist2 + np.where(below, (c[d] - lo[d]) ** 2, np.where(above, (c[d] - hi[d]) ** 2, 0.0))
        return np.nonzero(dist2 < r ** 2)[0]

    @staticmethod
    def _env(t, x, y, z, needs, lat_guard):
        """variables of an expression at points (x, y, z): radius, lat =
        asin(z / r) (0 at r = 0 where the caller guards it, else nan),
        lon = atan2(y, x)"""
        env = {"t": t, "x": x, "y": y, "z": z}
        if needs & {"radius", "lat", "lon"}:
            shape = np.broadcast(x, y, z).shape
            X, Y, Z = (np.broadcast_to(np.asarray(a, dtype=np.float64), shape) for a in (x, y, z))
            r = np.sqrt(X * X + Y * Y + Z * Z)
            env["radius"] = r
            if "lat" in needs:
                with np.errstate(all="ignore"):
                    lat = _asin(Z / r)
                env["lat"] = np.where(r == 0, 0.0, lat) if lat_guard else lat
            if "lon" in needs:
                env["lon"] = _atan2(Y, X).astype(np.float64) if shape else np.float64(math.atan2(Y, X))
        return env

    # --- initial state ------------------------------------------------------
    def _face_copies(self, value, t, cells_idx=None):
        """initialize_magnetic_field's integrate_face_B of every face copy:
        (2, 3, m) = [minus / plus copy][dimension][cell] of the normal
        component for the selected cells"""
        lo, hi, _ = self.emulated_bounds()
        if cells_idx is not None:
            lo, hi = lo[:, cells_idx], hi[:, cells_idx]
        needs = value.needs()
        m = lo.shape[1]

        def f(comp, x, y, z):
            shape = np.broadcast(x, y, z).shape
            return value(shape, **self._env(t, x, y, z, needs, True))[comp]

        out = np.empty((2, 3, m))
        for side in (0, 1):
            for dim in range(3):
                o = [a for a in range(3) if a != dim]     # the face's two directions
                pos = lo[dim] if side == 0 else hi[dim]
                s0, e0 = lo[o[0]], hi[o[0]]
                s1, e1 = lo[o[1]], hi[o[1]]
                d0, d1 = e0 - s0, e1 - s1

                # the cases of initialize.hpp:364-497 (per cell, by which
                # of the face's extents are zero)
                res = np.empty(m)
                cases = [(d0 == 0) & (d1 == 0), (d0 == 0) & (d1 != 0),
                         (d0 != 0) & (d1 == 0), (d0 != 0) & (d1 != 0)]
                for ci, sel in enumerate(cases):
                    if not sel.any():
                        continue
                    S0, E0, S1, E1, P = s0[sel], e0[sel], s1[sel], e1[sel], pos[sel]
                    D0, D1 = E0 - S0, E1 - S1

                    def pt(u, v, P=P):
                        c = [None, None, None]
                        c[dim] = P
                        c[o[0]] = u
                        c[o[1]] = v
                        return f(dim, *c)
                    if ci == 0:
                        res[sel] = pt(S0, S1)
                    elif ci == 1:
                        res[sel] = _gl7(lambda v: pt(S0, v), S1, E1) / D1
                    elif ci == 2:
                        res[sel] = _gl7(lambda u: pt(u, S1), S0, E0) / D0
                    else:
                        # the outer integral runs over the face's second
                        # direction, the inner one over its first
                        res[sel] = _gl7(lambda v: _gl7(lambda u: pt(u, v), S0, E0), S1, E1) / D0 / D1
                out[side, dim] = res
        return out

    def _background_faces(self):
        """Bg_B of the positive faces, (9, n): [3 * dim + component], the
        field averaged over each face as initialize_magnetic_field does
        (initialize.hpp:499-522)"""
        if self.background is None:
            return None
        lo, hi, _ = self.emulated_bounds()
        n = self.n
        out = np.empty((9, n))
        for dim in range(3):
            o = [a for a in range(3) if a != dim]
            pos = hi[dim]
            s0, e0, s1, e1 = lo[o[0]], hi[o[0]], lo[o[1]], hi[o[1]]
            d0, d1 = e0 - s0, e1 - s1
            res = np.empty((3, n))
            cases = [(d0 == 0) & (d1 == 0), (d0 == 0) & (d1 != 0),
                     (d0 != 0) & (d1 == 0), (d0 != 0) & (d1 != 0)]
            for ci, sel in enumerate(cases):
                if not sel.any():
                    continue
                S0, E0, S1, E1, Pd = s0[sel], e0[sel], s1[sel], e1[sel], pos[sel]

                def f(u, v, Pd=Pd):
                    c = [None, None, None]
                    c[dim] = Pd
                    c[o[0]] = u
                    c[o[1]] = v
                    return self.background_field(*c)
                if ci == 0:
                    res[:, sel] = f(S0, S1)
                elif ci == 1:
                    res[:, sel] = _gl7(lambda v: f(S0, v), S1, E1) / (E1 - S1)
                elif ci == 2:
                    res[:, sel] = _gl7(lambda u: f(u, S1), S0, E0) / (E0 - S0)
                else:
                    res[:, sel] = _gl7(lambda v: _gl7(lambda u: f(u, v), S0, E0), S1, E1) \
                        / (E0 - S0) / (E1 - S1)
            out[3 * dim:3 * dim + 3] = res
        return out

    def initial_state(self, t=None):
        """(state (8, n) face-once, background (9, n) or None) after the
        initialisation sequence of tests/mhd/test.cpp:420-450 (periodic
        configurations; bounded ones: bounded_initial_state)"""
        if self.bounded:
            raise ConfigError("a configuration with boundaries or closed dimensions: "
                              "use bounded_initial_state")
        t = self.time_start if t is None else t
        rho, mom, E, volB, F = self._initialise(t, None)
        gamma, mu0 = self.adiabatic_index, self.vacuum_permeability
        # apply_boundaries: update_vol_B with constant pressure (cells with
        # positive density); the volume field is unchanged here
        pos = rho > 0
        if pos.any():
            rr, mm, BB = rho[pos], mom[:, pos], volB[:, pos]
            kin = 0.5 / rr * (mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2])
            mag = 0.5 / mu0 * (BB[0] * BB[0] + BB[1] * BB[1] + BB[2] * BB[2])
            p_old = (E[pos] - kin - mag) * (gamma - 1)
            v = mm * (1 / rr)
            E[pos] = _energy(rr, v, p_old, BB, gamma, mu0)
        # sync_magnetic_field: the two copies of every face, the visit of
        # the cell with the smaller index first
        self._sync(F)
        state = np.empty((8, self.n))
        state[0] = rho
        state[1:4] = mom
        state[4] = E
        state[5:8] = F[0]
        return state, self._background_faces()

    def bounded_initial_state(self, t=None):
        """(state (14, n), background (9, n) or None, (vtype, ctype, finfo))
        of a configuration with boundaries: the cell and face types
        (set_solver_info, classify_faces), then initialize_magnetic_field,
        update_vol_B of the cells of type >= 0 and initialize_fluid
        (tests/mhd/test.cpp:412-444).  The state holds each cell's copies of
        its six faces; apply_boundaries and sync_magnetic_field come next
        (the run applies them on the device)."""
        t = self.time_start if t is None else t
        types = self.classify()
        rho, mom, E, volB, F = self._initialise(t, types[1])
        S = np.empty((14, self.n))
        S[0] = rho
        S[1:4] = mom
        S[4] = E
        S[5:8] = volB
        for d in range(3):
            S[8 + 2 * d] = F[0, d]
            S[9 + 2 * d] = F[1, d]
        return S, self._background_faces(), types

    @property
    def bounded(self):
        """boundaries or a closed dimension: the cell-typed step"""
        return not all(self.periodic) or any(
            v.value_boundaries or v.copy_boundaries for v in self._variables())

    def _variables(self):
        return (self.number_density, self.velocity, self.pressure, self.magnetic_field)

    def boundary_masks(self):
        """(in_value, in_copy), (4, n) each: cells overlapping a value / copy
        boundary geometry of each variable (test.cpp:394-399)"""
        iv = np.zeros((4, self.n), dtype=np.uint8)
        ic = np.zeros((4, self.n), dtype=np.uint8)
        for v, var in enumerate(self._variables()):
            for gid, _, _ in var.value_boundaries:
                iv[v, self.overlapping(gid)] = 1
            for gid in var.copy_boundaries:
                ic[v, self.overlapping(gid)] = 1
        return iv, ic

    def classify(self):
        """(vtype (4, n), ctype (n,), finfo (6, n)), see pamhd_amd.boundary"""
        from . import boundary
        iv, ic = self.boundary_masks()
        try:
            vtype, ctype, finfo = boundary.classify(self.cells, self.periodic, iv, ic)
        except boundary.ClassificationError as e:
            raise ConfigError(str(e)) from None
        # apply_magnetic_field_boundaries_staggered throws for a copy cell
        # whose two faces of one dimension are both boundary or both not
        # solved (source/mhd/boundaries.hpp:333-337)
        cp = vtype[3] == 4
        for d in range(3):
            lo, hi = finfo[2 * d], finfo[2 * d + 1]
            bad = cp & (((lo == 0) & (hi == 0)) | ((lo == -1) & (hi == -1)))
            if bad.any():
                c = int(np.nonzero(bad)[0][0])
                raise ConfigError("magnetic field copy cell %d has no normal face in dimension %d "
                                  "(the reference's boundary code rejects it)" % (c, d))
        return vtype, ctype, finfo

    def boundary_values(self, t):
        """The value boundaries at time t, resolved per cell (a later
        boundary overrides an earlier one on the cells they share):
        (mask (4, n) uint8, number density (n), velocity (3, n), pressure
        (n), face field (6, n)).  Fluid values at the cell centres, the
        field's normal component at the centres of the faces -x ... +z
        (apply_fluid_boundaries, source/mhd/boundaries.hpp:738-1026, and
        apply_magnetic_field_boundaries_staggered, :270-305, whose latitude
        is asin(z / x))."""
        from .boundary import nearest_stamp
        n = self.n
        mask = np.zeros((4, n), dtype=np.uint8)
        out = [np.zeros(n), np.zeros((3, n)), np.zeros(n), np.zeros((6, n))]
        _, _, ctr = self.cell_bounds()
        for v, var in enumerate(self._variables()):
            for gid, stamps, vals in var.value_boundaries:
                cells = self.overlapping(gid)
                if not len(cells):
                    continue
                val = vals[nearest_stamp(stamps, t)]
                c = ctr[:, cells]
                if v < 3:
                    env = self._env(t, c[0], c[1], c[2], val.needs(), False)
                    got = val((len(cells),), **env)
                    if v == 1:
                        out[1][:, cells] = got
                    else:
                        out[v][cells] = got[0]
                else:
                    for f in range(6):
                        d, side = f // 2, (1 if f % 2 else -1)
                        pos = c.copy()
                        pos[d] = pos[d] + side * self.cell_length[d] / 2
                        env = self._env_bfield(t, pos, val.needs())
                        out[3][f, cells] = val((len(cells),), **env)[d]
                mask[v, cells] = 1
        return (mask, *out)

    @staticmethod
    def _env_bfield(t, pos, needs):
        env = {"t": t, "x": pos[0], "y": pos[1], "z": pos[2]}
        if needs & {"radius", "lat", "lon"}:
            x, y, z = pos
            env["radius"] = np.sqrt(x * x + y * y + z * z)
            with np.errstate(all="ignore"):
                env["lat"] = _asin(z / x)
            env["lon"] = _atan2(y, x).astype(np.float64)
        return env

    def _initialise(self, t, ctype):
        """face copies, then the fluid: (rho, m, E, Vol_B, F) with F of
        shape (2, 3, n) = [minus / plus copy][dimension][cell]; Vol_B is
        the average of each cell's own copies where its type is >= 0 (all
        cells when ctype is None) and 0 elsewhere"""
        n = self.n
        gamma, mu0, mp = self.adiabatic_index, self.vacuum_permeability, self.proton_mass
        if gamma <= 1:
            raise ConfigError("adiabatic index must be > 1")
        # face B copies: defaults, then regions i with their own value
        F = self._face_copies(self.magnetic_field.default, t)
        for i, (gid, val) in enumerate(self.magnetic_field.regions):
            cells = self.overlapping(gid)
            if len(cells):
                F[:, :, cells] = self._face_copies(val, t, cells)
        # update_vol_B(constant pressure = false)
        volB = 0.5 * (F[0] + F[1])
        if ctype is not None:
            volB[:, ctype < 0] = 0.0
        # initialize_fluid: defaults at the emulated centres (lat guarded at r = 0)
        _, _, ctr = self.emulated_bounds()
        x, y, z = ctr
        shape = (n,)

        def at(var_value, sel=None, guard=True):
            xs, ys, zs = (x, y, z) if sel is None else (x[sel], y[sel], z[sel])
            env = self._env(t, xs, ys, zs, var_value.needs(), guard)
            return var_value(xs.shape, **env)

        rho = mp * at(self.number_density.default)[0]
        vel = at(self.velocity.default)
        p = at(self.pressure.default)[0]
        mom = vel * rho
        E = np.zeros(n)
        ok = (rho > 0) & (p > 0)
        E[ok] = _energy(rho[ok], vel[:, ok], p[ok], volB[:, ok], gamma, mu0)
        # regions: the value of region number geometry-id (the reference
        # indexes the fluid regions by the geometry id), lat unguarded
        nd = self.number_density.regions
        for gid, _v in nd:
            cells = self.overlapping(gid)
            if len(cells):
                rho[cells] = mp * at(_region(nd, gid, "number-density"), cells, False)[0]
        vr = self.velocity.regions
        for gid, _v in vr:
            cells = self.overlapping(gid)
            if len(cells):
                mom[:, cells] = rho[cells] * at(_region(vr, gid, "velocity"), cells, False)
        pr = self.pressure.regions
        for gid, _v in pr:
            cells = self.overlapping(gid)
            if len(cells):
                pv = at(_region(pr, gid, "pressure"), cells, False)[0]
                r = rho[cells]
                good = (r > 0) & (pv > 0)
                e = np.zeros(len(cells))
                if good.any():
                    rg = r[good]
                    v = mom[:, cells][:, good] * (1 / rg)
                    e[good] = _energy(rg, v, pv[good], volB[:, cells][:, good], gamma, mu0)
                E[cells] = e
        return rho, mom, E, volB, F

    def _sync(self, F):
        """sync_magnetic_field on the copies F[0] (minus faces) and F[1]
        (plus faces): each face is visited from both of its cells in cell
        order, each visit moving both copies half their difference toward
        each other.  (In a one-cell dimension both copies lie at the same
        point and are equal.)"""
        nx, ny, nz = self.cells
        idx = np.arange(self.n).reshape(nz, ny, nx)
        me = idx.reshape(-1)
        for d in range(3):
            up = np.roll(idx, -1, axis=2 - d).reshape(-1)     # +d neighbour
            P, M = F[1, d], F[0, d]
            for mask, first_from_me in ((me < up, True), (me > up, False)):
                c, u = me[mask], up[mask]
                p, m = P[c], M[u]
                for from_me in (first_from_me, not first_from_me):
                    if from_me:
                        diff = p - m
                        p = p - diff / 2
                        m = m + diff / 2
                    else:
                        diff = m - p
                        m = m - diff / 2
                        p = p + diff / 2
                P[c] = p
                M[u] = m

    # --- substepping options --------------------------------------------------
    def substep_options(self, t):
        """(substep-min, substep-max) at time t: ints, or per-cell int arrays
        of the expressions at the cell centres
        (set_minmax_substepping_period(time, ...), substepping.hpp:197-255)"""
        lo, hi = self.substep_min, self.substep_max
        if isinstance(lo, int) and isinstance(hi, int):
            return lo, hi
        _, _, ctr = self.cell_bounds()
        out = []
        for v in (lo, hi):
            if isinstance(v, int):
                out.append(np.full(self.n, v, dtype=np.int32))
                continue
            env = self._env(t, ctr[0], ctr[1], ctr[2], v.variables(), False)
            val = np.broadcast_to(np.asarray(v.evaluate(**env), dtype=np.float64), (self.n,))
            if not np.all(np.isfinite(val)):
                raise ConfigError("%r is not finite in every cell" % v.text)
            out.append(np.trunc(val).astype(np.int32))
        return out[0], out[1]

    def substeps_depend_on_time(self):
        return any(isinstance(v, Expression) and "t" in v.variables()
                   for v in (self.substep_min, self.substep_max))

    def boundaries_depend_on_time(self):
        """a value boundary with more than one time stamp or an expression
        of t"""
        for var in self._variables():
            for _, stamps, vals in var.value_boundaries:
                if len(stamps) > 1 or any("t" in v.needs() for v in vals):
                    return True
        return False

    def make_grid(self, precision="fp64"):
        """a solver.Grid holding the initial state (with the background
        field and the substep options at time-start); a grid with
        boundaries has them applied and its faces synced
        (test.cpp:446-457)"""
        from .solver import Grid
        if self.bounded:
            if precision != "fp64":
                raise ConfigError("grids with boundaries run in FP64")
            state, bg, types = self.bounded_initial_state()
            g = Grid(self.cells, self.cell_length, self.adiabatic_index,
                     self.vacuum_permeability, self.proton_mass, start=self.start,
                     periodic=self.periodic, types=types)
        else:
            state, bg = self.initial_state()
            g = Grid(self.cells, self.cell_length, self.adiabatic_index,
                     self.vacuum_permeability, self.proton_mass, precision=precision,
                     start=self.start)
        try:
            g.upload(state)
            if bg is not None:
                g.set_background_faces(bg)
            if precision == "fp64":
                lo, hi = self.substep_options(self.time_start)
                g.set_substepping(lo, hi)
            if self.bounded:
                g.set_boundary_values(*self.boundary_values(self.time_start))
                g.apply_boundaries(self.time_start)
        except Exception:
            g.close()
            raise
        return g


def _region(regions, gid, name):
    if gid >= len(regions):
        raise ConfigError("%s: geometry-id %d has no region of that index (the reference "
                          "reads the value of region number geometry-id)" % (name, gid))
    return regions[gid][1]


def _energy(rho, v, p, B, gamma, mu0):
    """get_total_energy_density, source/mhd/common.hpp:251-311"""
    kin = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    mag = 0.5 / mu0 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2])
    return p / (gamma - 1) + kin + mag


def load(path_or_obj):
    """Config from a JSON file path, a JSON string or a parsed object"""
    if isinstance(path_or_obj, dict):
        return Config(path_or_obj)
    if isinstance(path_or_obj, str) and path_or_obj.lstrip().startswith("{"):
        return Config(json.loads(path_or_obj))
    with open(path_or_obj, encoding="utf-8") as f:
        return Config(json.load(f), path=os.fspath(path_or_obj))


def run(cfg, precision="fp64", output_directory=None, max_steps=1 << 62, grid=None):
    """The program loop of tests/mhd/test.cpp:403-600 on the GPU: the final
    init (a zero-length step), then steps of timestep + apply_boundaries
    until time-start + time-length, never stepping over the end; the
    substep options are set again before every step when they depend on
    time.  With output_directory, writes mhd_<step>.dc files at the steps
    the reference saves (save-mhd-n).  Returns a dict of steps, end time,
    the final state and the largest substep period of the last step."""
    from . import dc
    own = grid is None
    g = cfg.make_grid(precision) if own else grid
    try:
        t = cfg.time_start
        t_end = cfg.time_start + cfg.time_length
        saves = []

        def save(step, time):
            if output_directory is None:
                return
            os.makedirs(output_directory, exist_ok=True)
            path = os.path.join(output_directory, "mhd_%09d.dc" % step)
            dc.write(path, g.download(), cfg.cells, cfg.start, cfg.cell_length, cfg.periodic,
                     sim_step=step, sim_time=time, adiabatic_index=cfg.adiabatic_index,
                     proton_mass=cfg.proton_mass, vacuum_permeability=cfg.vacuum_permeability)
            saves.append(path)

        per_step_options = precision == "fp64" and cfg.substeps_depend_on_time()
        timed_boundaries = cfg.bounded and cfg.boundaries_depend_on_time()
        # the state already holds apply_boundaries' refit of the
        # initialisation, so the loop starts at the final init (pamhd_run
        # would refit once more)
        g.timestep(cfg.solver, 0.0, cfg.time_step_factor)
        step = 0
        if cfg.save_n >= 0:
            save(0, t)
        next_save = cfg.save_n
        while t < t_end and step < max_steps:
            step += 1
            if per_step_options:
                g.set_substepping(*cfg.substep_options(t))
            if timed_boundaries:
                # the boundaries take their values at the step's end time
                dt = g.timestep(cfg.solver, t_end - t, cfg.time_step_factor)
                t += dt
                g.set_boundary_values(*cfg.boundary_values(t))
                g.apply_boundaries(t)
            else:
                dt = g.step(cfg.solver, t_end - t, cfg.time_step_factor)
                t += dt
            if (cfg.save_n >= 0 and t >= t_end) or (cfg.save_n > 0 and t >= next_save):
                if cfg.save_n > 0 and next_save <= t:
                    next_save += cfg.save_n * math.ceil(max(cfg.save_n, t - next_save) / cfg.save_n)
                save(step, t)
        return dict(steps=step, time=t, state=g.download(), saves=saves,
                    max_substep=g.last_max_substep())
    finally:
        if own:
            g.close()
