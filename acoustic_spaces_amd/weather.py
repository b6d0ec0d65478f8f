"""Outdoor weather as a speed of sound per cell: the effective-sound-speed field that the
scaled stepping kernels read (the layout's ``scale``, each cell's speed squared).

The reference's ``_Weather`` (engine/wave.py:457-627) over its scene model's
``Atmosphere`` (scene/model.py:3270-3480) gives each cell of an outdoor grid the still air's
speed of sound at its height above the ground, plus the wind's speed at that height times
the cosine between the wind's direction and the horizontal from the source (0 straight
above it): sound carried by the wind goes faster downwind and slower upwind, which bends
it. Over the grid, each cell's ratio to the grid's speed, at most 1 (``_Weather.ratios``).

Here the same, for one atmosphere over the whole scene:

- temperature T(z) = T_ground + gradient z + log_k ln(1 + z / roughness), or a measured
  table [height, temperature, wind] interpolated linearly (as at its top above it);
- wind w(z): none, constant, linear (speed z / height) or logarithmic
  (speed ln(1 + z / roughness) / ln(1 + height / roughness)), or the table's;
- the still air's speed c(z) = 331.3 sqrt(1 + T(z) / 273.15) (air.py:23);
- heights below the ground count as at the ground.

Weather zones (engine/weather.py, scene/model.py:2955-3067): parts of the ground (polygons
in plan, clipped to the ground's extent) with weather of their own, each given over the
atmosphere's (what a zone gives replaces the atmosphere's, what it leaves out stays). A cell
takes the weather of the first zone whose polygon holds its centre's (x, y) (even-odd rule,
edges half-open along y), else the atmosphere's own; the cosine is to that weather's wind.
A zone may start from a named preset weather. The speeds are formed plane by plane in double and
rounded once to float, as the reference's are.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

SOUND_SPEED_0C = 331.3  # m/s in still air at 0 degrees C (air.py:23)
KELVIN = 273.15

# The weathers a zone can start from whole (scene/model.py:2910-2922): a moderate wind, 5 m/s
# at its height; a sunny day's lapse (the air cooling with height); a clear night's
# inversion; still air.
PRESETS = {
    "neutral": {},
    "windy": {"wind_speed_m_s": 5.0},
    "sunny_day": {"temperature_log_k": -0.5, "temperature_gradient_k_m": -0.0098},
    "clear_night": {"temperature_log_k": 0.5, "temperature_gradient_k_m": 0.0},
}


@dataclasses.dataclass
class WeatherZone:
    """A part of the ground over which the weather differs (scene/model.py:3051-3067): the
    ``polygon``'s (x, y) corners in order (m); each field given (not None) replaces the
    atmosphere's: the wind's speed, direction, height and profile, the temperature's
    gradient and logarithmic term, the roughness, a measured table, and the temperature at
    the ground (C). Giving the gradient or the log term drops the atmosphere's table; giving
    a table drops them (scene/model.py:3377-3413). A ``preset`` (``PRESETS``) gives that
    weather whole first, its wind blowing the atmosphere's way at its height."""

    polygon: tuple
    wind_speed_m_s: float | None = None
    wind_direction_deg: float | None = None
    wind_height_m: float | None = None
    wind_profile: str | None = None
    temperature_gradient_k_m: float | None = None
    temperature_log_k: float | None = None
    roughness_m: float | None = None
    table: tuple | None = None
    temperature_c: float | None = None
    name: str | None = None
    preset: str | None = None

    def __post_init__(self):
        if self.preset is not None and self.preset not in PRESETS:
            raise ValueError(f"no atmosphere preset named {self.preset!r} (the presets are {', '.join(PRESETS)})")
        corners = np.asarray(self.polygon, float)
        if corners.ndim != 2 or corners.shape[1] != 2 or len(corners) < 3:
            raise ValueError("a weather zone's polygon is three (x, y) corners or more")
        if self.table is not None and (self.temperature_gradient_k_m is not None or self.temperature_log_k is not None):
            raise ValueError("give a weather zone a table or the temperature's gradient and log terms, not both")


@dataclasses.dataclass
class Atmosphere:
    """The weather over an outdoor scene, by height above the ground (scene/model.py:
    3270-3480): the temperature's gradient (K/m) and logarithmic coefficient (K), the
    ground's roughness length (m), the wind's speed (m/s) at ``wind_height_m``, its
    direction (degrees from +x towards +y), its profile (``log``, ``linear`` or
    ``constant``), or a measured ``table`` of (height m, temperature C, wind m/s) rows."""

    temperature_gradient_k_m: float = 0.0
    temperature_log_k: float = 0.0
    roughness_m: float = 0.05
    wind_speed_m_s: float = 0.0
    wind_direction_deg: float = 0.0
    wind_height_m: float = 10.0
    wind_profile: str = "log"
    table: tuple | None = None
    zones: tuple = ()

    def __post_init__(self):
        if self.wind_profile not in ("log", "linear", "constant"):
            raise ValueError(f"unknown wind profile {self.wind_profile!r}; use log, linear or constant")
        if not self.roughness_m > 0 or not self.wind_height_m > 0:
            raise ValueError("the roughness and the wind's height must be positive")
        if self.table is not None:
            rows = np.asarray(self.table, float)
            if rows.ndim != 2 or rows.shape[1] != 3 or len(rows) < 1 or np.any(np.diff(rows[:, 0]) <= 0):
                raise ValueError("the table holds (height, temperature, wind) rows by rising height")

    @property
    def still(self) -> bool:
        """No wind and one temperature: sound goes straight, the grid needs no field (with
        zones: in every weather, scene/model.py:3471-3480)."""
        if self.zones:
            return all(w.still for w in self.weathers())
        if self.table is not None:
            rows = np.asarray(self.table, float)
            return bool(np.all(rows[:, 1] == rows[0, 1]) and np.all(rows[:, 2] == 0.0))
        return (self.wind_speed_m_s == 0.0 and self.temperature_gradient_k_m == 0.0
                and self.temperature_log_k == 0.0)

    def weathers(self) -> list:
        """The atmosphere's own weather (first, without zones) and each zone's, resolved
        (scene/model.py:3354-3413)."""
        own = dataclasses.replace(self, zones=())
        return [own, *(self.resolved(z) for z in self.zones)]

    def resolved(self, zone: WeatherZone) -> "Atmosphere":
        """The atmosphere with ``zone``'s weather given over it."""
        values = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
        values["zones"] = ()
        if zone.preset is not None:
            # The preset's weather whole: no wind unless it has one (then blowing the
            # atmosphere's way, at its height, by its profile), no gradient terms, no table.
            preset = PRESETS[zone.preset]
            values.update(wind_speed_m_s=0.0, temperature_gradient_k_m=0.0, temperature_log_k=0.0, table=None)
            values.update(preset)
            if "wind_speed_m_s" not in preset and any(
                    getattr(zone, f) is not None for f in ("wind_speed_m_s", "wind_direction_deg", "wind_height_m",
                                                           "wind_profile")):
                # A zone's wind over a preset without one starts from the atmosphere's wind.
                values["wind_speed_m_s"] = self.wind_speed_m_s
        for name in ("wind_speed_m_s", "wind_direction_deg", "wind_height_m", "wind_profile",
                     "temperature_gradient_k_m", "temperature_log_k", "roughness_m"):
            if getattr(zone, name) is not None:
                values[name] = getattr(zone, name)
        if zone.table is not None:
            values["table"] = zone.table
        elif zone.temperature_gradient_k_m is not None or zone.temperature_log_k is not None:
            values["table"] = None
        if values["table"] is not None:
            values["temperature_gradient_k_m"] = values["temperature_log_k"] = 0.0
        return Atmosphere(**values)

    def ground_temperatures(self, ground_c: float) -> list:
        """The temperature at the ground (C) under each weather, in :meth:`weathers` order."""
        return [float(ground_c), *(float(ground_c) if z.temperature_c is None else float(z.temperature_c)
                                   for z in self.zones)]

    def direction(self) -> np.ndarray:
        """The unit vector (2,) the wind blows along."""
        a = math.radians(self.wind_direction_deg)
        return np.array([math.cos(a), math.sin(a)])

    def temperature(self, height, ground_c: float) -> np.ndarray:
        """The temperature (C) at ``height`` (m above the ground)."""
        z = np.maximum(np.asarray(height, float), 0.0)
        if self.table is not None:
            rows = np.asarray(self.table, float)
            return np.interp(z, rows[:, 0], rows[:, 1])
        return ground_c + self.temperature_gradient_k_m * z + self.temperature_log_k * np.log1p(z / self.roughness_m)

    def wind(self, height) -> np.ndarray:
        """The wind's speed (m/s) at ``height`` (m above the ground)."""
        z = np.maximum(np.asarray(height, float), 0.0)
        if self.table is not None:
            rows = np.asarray(self.table, float)
            return np.interp(z, rows[:, 0], rows[:, 2])
        v = self.wind_speed_m_s
        if v == 0.0:
            return np.zeros_like(z)
        if self.wind_profile == "constant":
            return np.full_like(z, v)
        if self.wind_profile == "linear":
            return v * z / self.wind_height_m
        return v * np.log1p(z / self.roughness_m) / math.log1p(self.wind_height_m / self.roughness_m)

    def sound_speed(self, height, ground_c: float) -> np.ndarray:
        """The still air's speed of sound (m/s) at ``height``."""
        return SOUND_SPEED_0C * np.sqrt(1.0 + self.temperature(height, ground_c) / KELVIN)


def inside(points, polygon) -> np.ndarray:
    """(n,) whether each of ``points`` (n, 2) lies inside ``polygon`` (m, 2), by the
    even-odd rule: the ray from the point towards +x crosses an odd number of its edges. An
    edge counts where exactly one of its ends lies strictly above the point's y and it
    passes to the right of the point at that y."""
    p = np.asarray(points, float).reshape(-1, 2)
    a = np.asarray(polygon, float)
    b = np.roll(a, -1, axis=0)
    y = p[:, 1][:, None]
    spans = (a[None, :, 1] > y) != (b[None, :, 1] > y)
    with np.errstate(divide="ignore", invalid="ignore"):
        x_at = a[None, :, 0] + (y - a[None, :, 1]) * (b[None, :, 0] - a[None, :, 0]) / (b[None, :, 1] - a[None, :, 1])
    return np.count_nonzero(spans & (p[:, 0][:, None] < x_at), axis=1) % 2 == 1


def clip(polygon, extent):
    """The part of ``polygon`` inside the rectangle ``extent`` (x_min, y_min, x_max,
    y_max), clipped against its four sides in turn (x_min, x_max, y_min, y_max); None where
    nothing with an area is left."""
    x0, y0, x1, y1 = (float(v) for v in extent)
    pts = [tuple(float(c) for c in q) for q in np.asarray(polygon, float)]
    for axis, bound, keep_above in ((0, x0, True), (0, x1, False), (1, y0, True), (1, y1, False)):
        keep = [(q[axis] >= bound) if keep_above else (q[axis] <= bound) for q in pts]
        out = []
        for k, q in enumerate(pts):
            prev, prev_keep = pts[k - 1], keep[k - 1]
            if keep[k] != prev_keep:
                t = (bound - prev[axis]) / (q[axis] - prev[axis])
                cut = [prev[0] + t * (q[0] - prev[0]), prev[1] + t * (q[1] - prev[1])]
                cut[axis] = bound
                out.append(tuple(cut))
            if keep[k]:
                out.append(q)
        pts = out
        if not pts:
            return None
    c = np.array(pts)
    tolerance = 1e-9 * max(x1 - x0, y1 - y0)
    c = c[np.linalg.norm(c - np.roll(c, 1, axis=0), axis=1) > tolerance]
    if len(c) < 3:
        return None
    area = 0.5 * float(np.sum(c[:, 0] * np.roll(c[:, 1], -1) - np.roll(c[:, 0], -1) * c[:, 1]))
    if abs(area) <= tolerance * max(x1 - x0, y1 - y0):
        return None
    return c


class Weather:
    """The effective speed of sound over a grid from a source (``_Weather``,
    engine/wave.py:457-627): the atmosphere over a ground at height ``ground`` (m) whose
    temperature is ``ground_c``, seen from ``source`` (3,); with zones, each over the part
    of the ground its polygon covers, clipped to ``extent`` (x_min, y_min, x_max, y_max:
    the ground's; None keeps the polygons whole)."""

    def __init__(self, atmosphere: Atmosphere, ground_c: float, ground: float, source, extent=None):
        self.atmosphere = atmosphere
        self.ground_c = float(ground_c)
        self.ground = float(ground)
        self.source = np.asarray(source, float).reshape(3)
        self.weathers = atmosphere.weathers()
        self.grounds = atmosphere.ground_temperatures(ground_c)
        self.polygons = []
        for zone in atmosphere.zones:
            poly = np.asarray(zone.polygon, float)
            self.polygons.append(poly if extent is None else clip(poly, extent))

    @property
    def zoned(self) -> bool:
        return bool(self.polygons)

    def zone_at(self, points) -> np.ndarray:
        """(n,) the weather over each of ``points`` (n, 2 or 3): 0 the atmosphere's own,
        k the k-th zone's (the first where zones overlap)."""
        p = np.asarray(points, float)
        p = p.reshape(-1, p.shape[-1])[:, :2]
        out = np.zeros(len(p), np.int64)
        for k, poly in enumerate(self.polygons, start=1):
            if poly is not None:
                out[(out == 0) & inside(p, poly)] = k
        return out

    def profiles(self, z, k: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """The still air's speed and the wind's at heights z (m, the scene's coordinates),
        under weather k (0: the atmosphere's own)."""
        height = np.asarray(z, float) - self.ground
        w = self.weathers[k]
        return w.sound_speed(height, self.grounds[k]), w.wind(height)

    def cosine(self, x, y, k: int = 0) -> np.ndarray:
        """The cosine between weather k's wind and the horizontal from the source to
        (x, y); 0 straight above or below the source."""
        wx, wy = self.weathers[k].direction()
        dx, dy = np.asarray(x, float) - self.source[0], np.asarray(y, float) - self.source[1]
        reach = np.hypot(dx, dy)
        along = wx * dx + wy * dy
        return np.divide(along, reach, out=np.zeros_like(along), where=reach > 0)

    def at(self, points) -> np.ndarray:
        """The effective speed of sound (m/s) at ``points`` (n, 3)."""
        p = np.asarray(points, float).reshape(-1, 3)
        zone = self.zone_at(p)
        out = np.zeros(len(p))
        for k in np.unique(zone).tolist():
            rows = zone == k
            still, wind = self.profiles(p[rows, 2], k)
            out[rows] = still + wind * self.cosine(p[rows, 0], p[rows, 1], k)
        return out

    def _grid(self, origin, cell: float, shape):
        """Each weather's profiles over the planes (K, nz) and cosines over a plane
        (K, ny, nx), and the weather over each column (ny, nx)."""
        origin = np.asarray(origin, float)
        nx, ny, nz = (int(v) for v in shape)
        xs = origin[0] + (np.arange(nx) + 0.5) * cell
        ys = origin[1] + (np.arange(ny) + 0.5) * cell
        zs = origin[2] + (np.arange(nz) + 0.5) * cell
        K = len(self.weathers) if self.zoned else 1
        still = np.empty((K, nz))
        wind = np.empty((K, nz))
        cos = np.empty((K, ny, nx))
        for k in range(K):
            still[k], wind[k] = self.profiles(zs, k)
            cos[k] = self.cosine(xs[None, :], ys[:, None], k)
        if not self.zoned:
            return still, wind, cos, None
        x, y = np.meshgrid(xs, ys)
        zone = self.zone_at(np.column_stack([x.ravel(), y.ravel()])).reshape(ny, nx)
        return still, wind, cos, zone

    # Moving air: each cell keeps the still air's speed of sound, and the wind moves the air.

    def winds_over(self, top: float) -> bool:
        """Whether the wind blows anywhere from the ground up to ``top`` (m, the scene's
        coordinates), under any of the weathers (engine/wave.py:906-912: 512 heights)."""
        z = np.linspace(self.ground, top, 512)
        return any(bool(np.any(w.wind(z - self.ground) != 0.0)) for w in self.weathers)

    def still_grid(self, origin, cell: float, shape):
        """The still air's speed of sound and the wind's speed over the planes (K, nz), each
        weather's wind direction (K, 2) and the weather over each column (ny, nx), or None
        without zones (K = 1) (``_Weather._still_grid``, engine/wave.py:578-593)."""
        origin = np.asarray(origin, float)
        nx, ny, nz = (int(v) for v in shape)
        zs = origin[2] + (np.arange(nz) + 0.5) * cell
        K = len(self.weathers) if self.zoned else 1
        still, wind = np.empty((K, nz)), np.empty((K, nz))
        for k in range(K):
            still[k], wind[k] = self.profiles(zs, k)
        towards = np.array([self.weathers[k].direction() for k in range(K)])
        if not self.zoned:
            return still, wind, towards, None
        x, y = np.meshgrid(origin[0] + (np.arange(nx) + 0.5) * cell, origin[1] + (np.arange(ny) + 0.5) * cell)
        zone = self.zone_at(np.column_stack([x.ravel(), y.ravel()])).reshape(ny, nx)
        return still, wind, towards, zone

    def still_at(self, point) -> float:
        """The still air's speed of sound (m/s) at ``point`` (3,), under the weather over it."""
        p = np.asarray(point, float).reshape(3)
        k = int(self.zone_at(p[None, :])[0]) if self.zoned else 0
        return float(self.profiles(p[2:3], k)[0][0])

    def moving_extremes(self, origin, cell: float, shape) -> tuple[float, float, float]:
        """The slowest and fastest still air's speed of sound over the grid's cells and the
        wind's fastest speed (m/s), of the weathers over it (engine/wave.py:600-607)."""
        still, wind, _, zone = self.still_grid(origin, cell, shape)
        if zone is not None:
            used = np.unique(zone)
            still, wind = still[used], wind[used]
        return float(still.min()), float(still.max()), float(np.abs(wind).max())

    def moving_fields(self, origin, cell: float, shape, speed: float, shift: float):
        """(nx ny nz,) float32 each, x fastest: each cell's still air's speed of sound over
        ``speed`` (the grid's), at most 1, and the air's velocity along x and y in cells per
        step (the wind's speed times its direction's component times ``shift``, a step over a
        cell); each plane in double, then rounded (engine/wave.py:609-627)."""
        still, wind, towards, zone = self.still_grid(origin, cell, shape)
        nx, ny, nz = (int(v) for v in shape)
        speeds = np.empty((nz, ny, nx), np.float32)
        fx = np.empty((nz, ny, nx), np.float32)
        fy = np.empty((nz, ny, nx), np.float32)
        for z in range(nz):
            if zone is None:
                speeds[z] = min(still[0, z] / speed, 1.0)
                fx[z] = wind[0, z] * towards[0, 0] * shift
                fy[z] = wind[0, z] * towards[0, 1] * shift
            else:
                speeds[z] = np.minimum(still[zone, z] / speed, 1.0)
                fx[z] = wind[zone, z] * towards[zone, 0] * shift
                fy[z] = wind[zone, z] * towards[zone, 1] * shift
        return speeds.reshape(-1), fx.reshape(-1), fy.reshape(-1)

    def extremes(self, origin, cell: float, shape) -> tuple[float, float]:
        """The slowest and fastest effective speed over the cells of a grid (of the
        weathers over it)."""
        still, wind, cos, zone = self._grid(origin, cell, shape)
        low, high = np.inf, -np.inf
        for k in ([0] if zone is None else np.unique(zone).tolist()):
            c = cos[k] if zone is None else cos[k][zone == k]
            ends = wind[k][:, None] * np.array([c.min(), c.max()])[None, :]
            low = min(low, float((still[k] + ends.min(axis=1)).min()))
            high = max(high, float((still[k] + ends.max(axis=1)).max()))
        return low, high

    def ratios(self, origin, cell: float, shape, speed: float) -> np.ndarray:
        """(nx ny nz,) float32, x fastest: each cell's effective speed of sound over
        ``speed`` (the grid's: at least the fastest), at most 1 (``_Weather.ratios``). The
        cosines of a plane are formed once; each plane's speeds in double, then rounded."""
        still, wind, cos, zone = self._grid(origin, cell, shape)
        nz = still.shape[1]
        ny, nx = cos.shape[1:]
        out = np.empty((nz, ny, nx), np.float32)
        if zone is None:
            for z in range(nz):
                out[z] = np.minimum((still[0, z] + wind[0, z] * cos[0]) / speed, 1.0)
            return out.reshape(-1)
        iy, ix = np.indices(zone.shape)
        column_cos = cos[zone, iy, ix]  # each column's cosine to its own weather's wind
        for z in range(nz):
            out[z] = np.minimum((still[zone, z] + wind[zone, z] * column_cos) / speed, 1.0)
        return out.reshape(-1)
