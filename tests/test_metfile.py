"""Forcing preprocessing (solweig_project_amd.metfile) on the CPU: the met
file text format, ERA5 conversions against meteorological conventions and
closed forms, the nighttime UHI cycle against hand computation, and per-tile
sampling (local-day window, box mean, nearest-point fallback)."""
import datetime as dt

import numpy as np
import pytest

from solweig_project_amd import metfile as MF


def _row(**over):
    r = dict(zip(MF.COLUMNS, [2021, 180, 13, 0] + [-999.0] * 5 + [2.123456, 55.5, 28.25, 101.3, 0.0, 812.345]
                 + [-999.0] * 8 + [123.4, 1.25]))
    r.update(over)
    return [r[c] for c in MF.COLUMNS]


def test_row_format_matches_the_met_file_layout():
    line = MF.format_row(_row())
    assert line == ("2021 180 13 0 -999.00 -999.00 -999.00 -999.00 -999.00 2.12346 55.50 28.25 101.30 0.00 812.35 "
                    "-999.00 -999.00 -999.00 -999.00 -999.00 -999.00 -999.00 -999.00 123.40 1.25")
    with pytest.raises(ValueError):
        MF.format_row([1, 2, 3])


def test_write_read_round_trip(tmp_path):
    rows = [_row(it=h, Kdn=100.0 * h) for h in range(24)]
    p = tmp_path / "metfile_0_0_2021-06-29.txt"
    MF.write_metfile(p, rows)
    assert p.read_text().splitlines()[0] == " ".join(MF.COLUMNS)
    met = MF.read_metfile(p)
    assert met.shape == (24, 25)
    assert np.array_equal(met[:, 2], np.arange(24)) and np.array_equal(met[:, 14], 100.0 * np.arange(24))
    assert np.allclose(met[:, 9], 2.12346) and np.allclose(met[:, 24], 1.25)
    one = tmp_path / "one.txt"
    MF.write_metfile(one, rows[:1])
    assert MF.read_metfile(one).shape == (1, 25)
    short = tmp_path / "short.txt"
    short.write_text("a b\n1 2\n")
    with pytest.raises(ValueError, match="at least 23"):
        MF.read_metfile(short)


def test_read_metfile_feeds_the_day_loop(tmp_path):
    """A written met file drives the host chain like the in-memory table it came from."""
    from solweig_project_amd import met as M
    from solweig_project_amd import sunpos, synth
    sun = sunpos.day_sun_table(30.0, 0.0, 0, 2021, 180)
    met = synth.make_met(sun.altitude, seed=3)
    met[:, 9] = np.round(met[:, 9], 5)
    met[:, 10:15] = np.round(met[:, 10:15], 2)
    met[:, 23:25] = np.round(met[:, 23:25], 2)
    p = tmp_path / "m.txt"
    MF.write_metfile(p, met)
    back = MF.read_metfile(p)
    assert np.array_equal(back, met)
    a = M.DayLoop(back, sun, latitude=30.0, landcover=M.landcover_constants(False), scale=1.0, amax=10.0,
                  rows=8, cols=8)
    assert np.array_equal(a.wet_bulb.tw_c,
                          M.wet_bulb_from_rh(met[:, 12] * 1000, met[:, 11] + met[:, 24] + 273.15, met[:, 10]).tw_c)


def test_era5_conventions():
    z = np.zeros((4, 1, 1))
    u = np.array([0.0, 5.0, 0.0, -5.0]).reshape(4, 1, 1)
    v = np.array([-5.0, 0.0, 5.0, 0.0]).reshape(4, 1, 1)
    f = MF.era5_fields(z + 300.0, z + 300.0, z + 101325.0, u, v, z + 3600.0 * 500.0)
    # meteorological direction: the wind blows FROM north, west, south, east
    assert np.allclose(f["WDIR"].ravel(), [0.0, 270.0, 180.0, 90.0])
    assert np.allclose(f["WIND"], 5.0) and np.allclose(f["SWDOWN"], 500.0)
    assert np.allclose(f["RH2"], 100.0)                 # dewpoint = temperature
    calm = MF.era5_fields(z[:1] + 290, z[:1] + 280, z[:1], z[:1] + 0.001, z[:1], z[:1])
    assert calm["WDIR"].ravel()[0] == 0.0
    assert MF.magnus_svp_hpa(0.0) == pytest.approx(6.112)
    e = MF.magnus_svp_hpa(np.array([21.85, 26.85]))
    assert calm["RH2"].dtype == np.float32
    rh = MF.era5_fields(z[:1] + 300.0, z[:1] + 295.0, z[:1], z[:1], z[:1], z[:1])["RH2"].ravel()[0]
    assert rh == pytest.approx(100 * e[0] / e[1], rel=1e-6)
    assert MF.era5_fields(z[:1] + 280, z[:1] + 290, z[:1], z[:1], z[:1], z[:1])["RH2"].ravel()[0] == 100.0


def test_night_sine_profile():
    sw = np.array([0, 0, 0, 300, 600, 300, 4, 0, 100, 0.0])
    amp = np.full(10, 2.0)
    out = MF.night_sine_profile(sw, amp, 0, 0)
    assert np.allclose(out[:3], 2.0 * np.sin(np.linspace(0, np.pi, 3)))
    assert np.allclose(out[6:8], 2.0 * np.sin(np.linspace(0, np.pi, 2)))
    assert np.all(out[[3, 4, 5, 8, 9]] == 0.0)             # day, and a one-step night run
    ext = MF.night_sine_profile(sw, amp, 0, 1)             # a run extended by one step after it
    assert np.allclose(ext[:4], 2.0 * np.sin(np.linspace(0, np.pi, 4)))
    assert np.all(MF.night_sine_profile(sw, -amp, 0, 0) == 0.0)


def test_uhi_cycle_by_hand():
    times = np.arange("2000-08-13T00", "2000-08-15T00", dtype="datetime64[h]")
    h = np.arange(48) % 24
    sw = np.where((h >= 7) & (h <= 17), 800.0 * np.sin(np.pi * (h - 6) / 12.0), 0.0)     # night: h <= 6, h >= 18
    t2 = 295.0 + 8.0 * np.sin(np.pi * (h - 9) / 12.0)
    t2[24:] += 2.0 * np.sin(np.pi * (h[24:] - 9) / 12.0)       # a larger range on day 2
    u = np.full(48, 3.0)
    v = np.full(48, 4.0)
    got = MF.uhi_cycle(times, t2[:, None, None], sw[:, None, None], u[:, None, None], v[:, None, None])[:, 0, 0]
    amp = np.empty(48)
    for d in range(2):
        sl = slice(24 * d, 24 * d + 24)
        amp[sl] = ((sw[sl] / MF.RHO_CP).mean() * (t2[sl].max() - t2[sl].min()) ** 3 / 5.0) ** 0.25
    assert amp[30] > amp[0]
    # the night from 18 h to 6 h crosses midnight: its shape takes the mean of both days' amplitudes
    for night in (np.arange(0, 7), np.arange(18, 31), np.arange(42, 48)):
        exp = amp[night].mean() * np.sin(np.linspace(0, np.pi, night.size))
        assert np.allclose(got[night], exp, rtol=1e-6, atol=1e-6), night
    assert np.all(got[7:18] == 0) and np.all(got[31:42] == 0) and got.dtype == np.float32


def _grid():
    lat = np.array([29.0, 29.25, 29.5])
    lon = np.array([77.0, 77.25, 77.5, 77.75])
    lon2d, lat2d = np.meshgrid(lon, lat)
    times = np.arange("2000-08-12T12", "2000-08-14T12", dtype="datetime64[h]")
    T = times.size
    base = np.arange(12, dtype=float).reshape(3, 4)
    fields = {"T2": 290.0 + base[None] + np.arange(T)[:, None, None] * 0.01, "RH2": np.full((T, 3, 4), 50.0),
              "PSFC": np.full((T, 3, 4), 100000.0), "WIND": 1.0 + base[None].repeat(T, 0),
              "SWDOWN": np.zeros((T, 3, 4)), "WDIR": np.full((T, 3, 4), 370.0)}
    return fields, lat2d, lon2d, times


def test_sample_tile_box_mean_and_local_day():
    fields, lat2d, lon2d, times = _grid()
    # a box strictly around the 2 x 2 block of points lat 29.25-29.5, lon 77.25-77.5
    box = MF.TileBox(77.1, 29.1, 77.6, 29.6)
    rows = MF.sample_tile(fields, lat2d, lon2d, times, box, "Asia/Kolkata", "2000-08-13")
    # IST = UTC+5:30: local 00:00-23:59 on the 13th = 18:30 on the 12th to 18:29 on the 13th UTC
    assert rows.shape == (24, 25)
    assert np.array_equal(rows[:, 2], np.r_[np.arange(0, 24)]) and np.all(rows[:, 3] == 30)
    assert np.all(rows[:, 1] == 226) and np.all(rows[:, 0] == 2000)
    first = int(np.flatnonzero(times == np.datetime64("2000-08-12T19"))[0])
    inside = [5, 6, 9, 10]
    b = np.arange(12.0)
    assert np.allclose(rows[:, 11], 290.0 + b[inside].mean() + (first + np.arange(24)) * 0.01 - 273.15)
    assert np.allclose(rows[:, 9], 1.0 + b[inside].mean())
    assert np.allclose(rows[:, 12], 100.0) and np.allclose(rows[:, 23], 10.0)
    assert np.all(rows[:, 24] == 0.0) and np.all(rows[:, 13] == 0.0) and np.all(rows[:, 4:9] == -999.0)


def test_sample_tile_nearest_point_and_missing_values():
    fields, lat2d, lon2d, times = _grid()
    small = MF.TileBox(77.49, 29.24, 77.51, 29.26)       # much smaller than a 0.25 degree cell
    fields = dict(fields)
    del fields["RH2"]
    fields["UHI_CYCLE"] = np.full((times.size, 3, 4), 1.5)
    rows = MF.sample_tile(fields, lat2d, lon2d, times, small, "UTC", "2000-08-13")
    assert rows.shape == (24, 25)
    assert np.allclose(rows[:, 9], 1.0 + 6.0)            # grid point (29.25, 77.5)
    assert np.all(rows[:, 10] == -999.0)                 # no RH field
    assert np.allclose(rows[:, 24], 1.5)
    assert np.all(MF.sample_tile(fields, lat2d, lon2d, times, small, "UTC", "2000-08-13", use_uhi=False)[:, 24] == 0)
    with pytest.raises(ValueError, match="no forcing records"):
        MF.sample_tile(fields, lat2d, lon2d, times, small, "UTC", dt.date(2001, 1, 1))
