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

The reference's weather zones (polygons with weather of their own, engine/weather.py) are
not here. The speeds are formed plane by plane in double and rounded once to float, as the
reference's are.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

SOUND_SPEED_0C = 331.3  # m/s in still air at 0 degrees C (air.py:23)
KELVIN = 273.15


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
        """No wind and one temperature: sound goes straight, the grid needs no field."""
        if self.table is not None:
            rows = np.asarray(self.table, float)
            return bool(np.all(rows[:, 1] == rows[0, 1]) and np.all(rows[:, 2] == 0.0))
        return (self.wind_speed_m_s == 0.0 and self.temperature_gradient_k_m == 0.0
                and self.temperature_log_k == 0.0)

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


class Weather:
    """The effective speed of sound over a grid from a source (``_Weather``,
    engine/wave.py:457-627): the atmosphere over a ground at height ``ground`` (m) whose
    temperature is ``ground_c``, seen from ``source`` (3,)."""

    def __init__(self, atmosphere: Atmosphere, ground_c: float, ground: float, source):
        self.atmosphere = atmosphere
        self.ground_c = float(ground_c)
        self.ground = float(ground)
        self.source = np.asarray(source, float).reshape(3)

    def profiles(self, z) -> tuple[np.ndarray, np.ndarray]:
        """The still air's speed and the wind's at heights z (m, the scene's coordinates)."""
        height = np.asarray(z, float) - self.ground
        return self.atmosphere.sound_speed(height, self.ground_c), self.atmosphere.wind(height)

    def cosine(self, x, y) -> np.ndarray:
        """The cosine between the wind and the horizontal from the source to (x, y); 0
        straight above or below the source."""
        wx, wy = self.atmosphere.direction()
        dx, dy = np.asarray(x, float) - self.source[0], np.asarray(y, float) - self.source[1]
        reach = np.hypot(dx, dy)
        along = wx * dx + wy * dy
        return np.divide(along, reach, out=np.zeros_like(along), where=reach > 0)

    def at(self, points) -> np.ndarray:
        """The effective speed of sound (m/s) at ``points`` (n, 3)."""
        p = np.asarray(points, float).reshape(-1, 3)
        still, wind = self.profiles(p[:, 2])
        return still + wind * self.cosine(p[:, 0], p[:, 1])

    def extremes(self, origin, cell: float, shape) -> tuple[float, float]:
        """The slowest and fastest effective speed over the cells of a grid."""
        origin = np.asarray(origin, float)
        centres = [origin[a] + (np.arange(int(shape[a])) + 0.5) * cell for a in range(3)]
        still, wind = self.profiles(centres[2])
        cos = self.cosine(centres[0][None, :], centres[1][:, None])
        ends = wind[:, None] * np.array([cos.min(), cos.max()])[None, :]
        return float((still + ends.min(axis=1)).min()), float((still + ends.max(axis=1)).max())

    def ratios(self, origin, cell: float, shape, speed: float) -> np.ndarray:
        """(nx ny nz,) float32, x fastest: each cell's effective speed of sound over
        ``speed`` (the grid's: at least the fastest), at most 1 (``_Weather.ratios``). The
        cosines of a plane are formed once; each plane's speeds in double, then rounded."""
        origin = np.asarray(origin, float)
        nx, ny, nz = (int(v) for v in shape)
        xs = origin[0] + (np.arange(nx) + 0.5) * cell
        ys = origin[1] + (np.arange(ny) + 0.5) * cell
        zs = origin[2] + (np.arange(nz) + 0.5) * cell
        still, wind = self.profiles(zs)
        cos = self.cosine(xs[None, :], ys[:, None])  # (ny, nx)
        out = np.empty((nz, ny, nx), np.float32)
        for z in range(nz):
            out[z] = np.minimum((still[z] + wind[z] * cos) / speed, 1.0)
        return out.reshape(-1)
