"""Save-state reader/writer (app/io/ParticlesIO.java, app/io/MatrixIO.java,
app/PhysicsSettingsToml.java, app/Main.java:1258-1348): Java's Double.toString
layout, lossless round trips, the zip's entries and the loader's rules."""
from __future__ import annotations

import io
import struct
import zipfile

import numpy as np
import pytest

from particle_life import saves
from particle_life.saves import java_double


@pytest.mark.parametrize("x,s", [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1.0, "-1.0"), (0.02, "0.02"), (0.85, "0.85"),
    (0.001, "0.001"), (9.99e-4, "9.99E-4"), (1e7, "1.0E7"), (9999999.0, "9999999.0"),
    (1234567.0, "1234567.0"), (12345678.0, "1.2345678E7"), (100.0, "100.0"),
    (0.1 + 0.2, "0.30000000000000004"), (1 / 3, "0.3333333333333333"), (-2.5e-10, "-2.5E-10"),
    (1e-5, "1.0E-5"), (1e22, "1.0E22"), (2.0 ** -24, "5.960464477539063E-8"),
    (float("nan"), "NaN"), (float("inf"), "Infinity"), (float("-inf"), "-Infinity"),
])
def test_java_double_layout(x, s):
    # Double.toString's documented layout: plain in [1e-3, 1e7), else d.dddE<n>
    assert java_double(x) == s


def test_java_double_round_trips_bit_exactly():
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2 ** 63 - 2 ** 52, 20000, dtype=np.int64)     # finite positive doubles
    xs = np.concatenate([bits.view(np.float64), -bits[:100].view(np.float64),
                         rng.random(5000), rng.random(5000) * 1e-3, rng.random(1000) * 1e7])
    for x in xs:
        y = float(java_double(x))
        assert struct.pack("<d", y) == struct.pack("<d", float(x)), (x, java_double(x))


def test_particles_tsv_round_trip_and_header():
    rng = np.random.default_rng(4)
    pos = rng.random((50, 3))
    vel = rng.normal(size=(50, 3)) * 1e-3
    typ = rng.integers(0, 7, 50)
    text = saves.write_particles(pos, vel, typ)
    lines = text.splitlines()
    assert lines[0] == "x\ty\tvx\tvy\tcolor" and len(lines) == 51
    assert lines[1].split("\t") == [java_double(pos[0, 0]), java_double(pos[0, 1]),
                                    java_double(vel[0, 0]), java_double(vel[0, 1]), str(typ[0])]
    p2, v2, t2 = saves.read_particles(text)
    np.testing.assert_array_equal(p2[:, :2], pos[:, :2])
    np.testing.assert_array_equal(v2[:, :2], vel[:, :2])
    assert (p2[:, 2] == 0).all() and (v2[:, 2] == 0).all()        # loadParticles sets z = 0
    np.testing.assert_array_equal(t2, typ)


def test_matrix_tsv_round_trip():
    m = np.random.default_rng(5).uniform(-1, 1, (6, 6))
    text = saves.write_matrix(m)
    assert len(text.splitlines()) == 6 and text.splitlines()[0].count("\t") == 5
    np.testing.assert_array_equal(saves.read_matrix(text), m)


def test_physics_toml_keys_defaults_and_unknown():
    text = saves.write_physics_toml(False, 0.035, 0.7491535384383408, 2.0)
    assert 'boundaries = "clamped"' in text and "radius = 0.035" in text
    s = saves.read_physics_toml(text)
    assert s == {"wrap": False, "rmax": 0.035, "friction": 0.7491535384383408, "force": 2.0}
    # missing keys keep the current values (PhysicsSettingsToml defaults here)
    assert saves.read_physics_toml("radius = 0.05\n") == {"wrap": True, "rmax": 0.05,
                                                           "friction": 0.85, "force": 1.0}
    with pytest.raises(ValueError, match="dt"):
        saves.read_physics_toml("dt = 0.02\n")


def test_zip_round_trip_and_entries():
    rng = np.random.default_rng(6)
    pos, vel, typ = rng.random((20, 3)), rng.random((20, 3)), rng.integers(0, 4, 20)
    m = rng.uniform(-1, 1, (4, 4))
    buf = io.BytesIO()
    saves.save_state(buf, pos, vel, typ, m, True, 0.02, 0.85, 1.0)
    buf.seek(0)
    assert sorted(zipfile.ZipFile(buf).namelist()) == ["matrix.tsv", "particles.tsv", "physics.toml"]
    buf.seek(0)
    st = saves.load_state(buf)
    np.testing.assert_array_equal(st["particles"][0][:, :2], pos[:, :2])
    np.testing.assert_array_equal(st["matrix"], m)
    assert st["settings"]["rmax"] == 0.02
    # partial zips: only what is present; img.png ignored, other entries reported and skipped
    b2 = io.BytesIO()
    with zipfile.ZipFile(b2, "w") as z:
        z.writestr("matrix.tsv", saves.write_matrix(m))
        z.writestr("img.png", b"\x89PNG")
        z.writestr("notes.txt", "x")
    b2.seek(0)
    with pytest.warns(UserWarning, match="notes.txt"):
        st2 = saves.load_state(b2)
    assert set(st2) == {"matrix"}
