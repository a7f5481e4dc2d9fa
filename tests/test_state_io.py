"""State I/O: the restart stream (inirun.F90:598-700, ssave-cdf-instant.F90:
100-830) and dncoef.h1 (constant_mod.F90:331-364), host side, no GPU.

The reader is pinned against byte streams built here with `struct` from the
Fortran sequential-unformatted rule (4-byte record length, payload, length)
in the record order of inirun.F90, not by the library's own writer; the
writer must then reproduce those bytes exactly.  The reference builds its
Fortran with -convert big_endian (Makefile:4), so the streams are
big-endian: record lengths '>i', doubles '>f8', the date '>ii'; the native
little-endian form is an option."""
import os
import struct

import numpy as np
import pytest

import licom_synth as ls
import licom_tracer as lt


def fortran_records(payloads, order=">"):
    out = bytearray()
    for p in payloads:
        out += struct.pack(order + "i", len(p)) + p + struct.pack(order + "i", len(p))
    return bytes(out)


def restart_payloads(f, km, month, day, order=">"):
    """Records in the order inirun.F90:612-686 reads them."""
    dt = np.dtype(order + "f8")
    b = lambda a: np.ascontiguousarray(a).astype(dt).tobytes()  # noqa: E731
    recs = [b(f["h0"])]
    for a in (f["u"], f["v"], f["at"][0], f["at"][1], f["ws"]):
        recs += [b(a[k]) for k in range(km)]
    recs += [b(f[k]) for k in ("su", "sv", "swv", "sshf", "lthf", "fresh")]
    recs.append(struct.pack(order + "ii", month, day))
    return recs


def random_fields(nx, ny, km, seed=0):
    rng = np.random.default_rng(seed)
    f = {k: rng.normal(size=(ny, nx)) for k in lt.RESTART_2D}
    f.update(u=rng.normal(size=(km, ny, nx)), v=rng.normal(size=(km, ny, nx)),
             ws=rng.normal(size=(km, ny, nx)), at=rng.normal(size=(2, km, ny, nx)))
    return f


def test_reader_takes_fortran_records_in_inirun_order(tmp_path):
    nx, ny, km = 9, 6, 4
    f = random_fields(nx, ny, km)
    p = tmp_path / "fort.22"
    p.write_bytes(fortran_records(restart_payloads(f, km, 7, 23)))
    r = lt.read_restart(str(p), nx, ny, km)
    for k in f:
        np.testing.assert_array_equal(r[k], f[k], err_msg=k)
    assert (r["number_month"], r["number_day"]) == (7, 23)
    # the same records in the wrong byte order are refused (record length)
    with pytest.raises(lt.LicomError, match="record 0 holds"):
        lt.read_restart(str(p), nx, ny, km, byte_order="little")


def test_little_endian_option(tmp_path):
    nx, ny, km = 7, 5, 3
    f = random_fields(nx, ny, km, seed=6)
    p = tmp_path / "le"
    p.write_bytes(fortran_records(restart_payloads(f, km, 2, 28, "<"), "<"))
    r = lt.read_restart(str(p), nx, ny, km, byte_order="little")
    for k in f:
        np.testing.assert_array_equal(r[k], f[k], err_msg=k)
    assert (r["number_month"], r["number_day"]) == (2, 28)
    kmt = np.full((ny, nx), km, np.int32)
    q = tmp_path / "le2"
    lt.write_restart(str(q), f, kmt, kmt, 2, 28, byte_order="little")
    assert q.read_bytes() == p.read_bytes()
    with pytest.raises(lt.LicomError):
        lt.read_restart(str(p), nx, ny, km, byte_order="middle")


def test_writer_reproduces_the_record_stream_and_masks_land(tmp_path):
    nx, ny, km = 9, 6, 4
    rng = np.random.default_rng(1)
    f = random_fields(nx, ny, km, seed=2)
    kmt = rng.integers(0, km + 1, (ny, nx)).astype(np.int32)
    kmu = rng.integers(0, km + 1, (ny, nx)).astype(np.int32)
    before = {k: v.copy() for k, v in f.items()}
    p = tmp_path / "restart"
    lt.write_restart(str(p), f, kmt, kmu, 3, 9)
    for k in f:                                   # inputs untouched
        np.testing.assert_array_equal(f[k], before[k])
    # the expected stream: land zeroed as ssave-cdf-instant.F90:173-191 does
    lev = np.arange(km)[:, None, None]
    vit, viv = lev < kmt[None], lev < kmu[None]
    g = dict(f)
    g["h0"] = np.where(kmt > 0, f["h0"], 0.0)
    g["u"], g["v"] = np.where(viv, f["u"], 0.0), np.where(viv, f["v"], 0.0)
    g["ws"] = np.where(vit, f["ws"], 0.0)
    g["at"] = np.where(vit[None], f["at"], 0.0)
    assert p.read_bytes() == fortran_records(restart_payloads(g, km, 3, 9))
    r = lt.read_restart(str(p), nx, ny, km)
    for k in g:
        np.testing.assert_array_equal(r[k], g[k], err_msg=k)


def test_reader_rejects_bad_streams(tmp_path):
    nx, ny, km = 5, 4, 2
    good = fortran_records(restart_payloads(random_fields(nx, ny, km), km, 1, 1))
    p = tmp_path / "r"
    p.write_bytes(good[:-20])                                  # truncated
    with pytest.raises(lt.LicomError, match="record"):
        lt.read_restart(str(p), nx, ny, km)
    p.write_bytes(good)
    with pytest.raises(lt.LicomError, match="record 0 holds"):  # other grid
        lt.read_restart(str(p), nx + 1, ny, km)
    bad = bytearray(good)
    bad[4 + nx * ny * 8] ^= 0x01                               # first trailing length
    p.write_bytes(bytes(bad))
    with pytest.raises(lt.LicomError, match="trailing"):
        lt.read_restart(str(p), nx, ny, km)
    with pytest.raises(lt.LicomError):
        lt.read_restart(str(tmp_path / "missing"), nx, ny, km)


def test_eos_table_list_directed_and_scaled(tmp_path):
    km = 5
    rng = np.random.default_rng(4)
    to, so, po = rng.uniform(-2, 20, km), rng.uniform(-1e-3, 1e-3, km), rng.uniform(20, 30, km)
    C = rng.normal(size=(km, 9)) * 1e-4
    fmt = lambda x: f"{x:.17E}".replace("E", "D")  # noqa: E731  Fortran exponent letter
    lines = ["  reference temperature to", " ".join(fmt(x) for x in to[:3]),
             ",".join(fmt(x) for x in to[3:]),
             "  reference salinity so", "  ".join(f"{x:.17e}" for x in so),
             "  reference density po", f"{km}*{float(po[0])!r}"]
    for k in range(km):
        lines.append(f"  coefficients of level {k + 1}")
        lines.append(" ".join(fmt(x) for x in C[k, :5]))
        lines.append(", ".join(fmt(x) for x in C[k, 5:]))
    p = tmp_path / "dncoef.h1"
    p.write_text("\n".join(lines) + "\n")
    t, s, pp, e = lt.read_eos(str(p), km)
    np.testing.assert_array_equal(t, to)
    np.testing.assert_array_equal(s, so)
    np.testing.assert_array_equal(pp, np.full(km, po[0]))        # repeat count r*c
    # eosc[m][k] = 1000 C(k,m) (constant_mod.F90:357-363), licom_grid.eosc
    np.testing.assert_array_equal(e, (C * 1000.0).T)
    p.write_text("\n".join(lines[:8]) + "\n")                    # ends in level 1
    with pytest.raises(lt.LicomError, match="C\\(k,1:9\\)"):
        lt.read_eos(str(p), km)


def test_restart_round_trip_restores_the_engine_state(tmp_path):
    """A synthetic state written as a restart and read back into the engine's
    layout (two ghost cells, cyclic i, closed rows) gives back the same T/S
    on every wet cell and 0 on land."""
    c = ls.make_case("C1", nx=24, ny=16, km=6, depth_min=-2000.0, sigma=1.5, seed=5)
    kmt = c.grid["kmt"]
    kmt_g = np.ascontiguousarray(kmt[2:-2, 2:-2])
    kmu_g = kmt_g.copy()
    at = c.state["at"]
    f = {k: np.zeros((c.ny, c.nx)) for k in lt.RESTART_2D}
    f.update(u=np.zeros((c.km, c.ny, c.nx)), v=np.zeros((c.km, c.ny, c.nx)),
             ws=np.zeros((c.km, c.ny, c.nx)), at=np.ascontiguousarray(at[:, :, 2:-2, 2:-2]))
    f["h0"] = np.ascontiguousarray(c.state["h0f"][2:-2, 2:-2])
    p = tmp_path / "restart"
    lt.write_restart(str(p), f, kmt_g, kmu_g)
    r = lt.read_restart(str(p), c.nx, c.ny, c.km)
    back = lt.global_to_halo(r["at"])
    wet = np.arange(c.km)[:, None, None] < kmt[None]
    for n in range(2):
        np.testing.assert_array_equal(back[n][:, 2:-2, 2:-2][wet[:, 2:-2, 2:-2]],
                                      at[n][:, 2:-2, 2:-2][wet[:, 2:-2, 2:-2]])
        assert not np.any(back[n][:, 2:-2, 2:-2][~wet[:, 2:-2, 2:-2]])
        # ghost columns are the cyclic copies, ghost rows 0
        np.testing.assert_array_equal(back[n][:, 2:-2, :2], back[n][:, 2:-2, -4:-2])
        assert not np.any(back[n][:, :2]) and not np.any(back[n][:, -2:])
    assert os.path.getsize(p) == (1 + 5 * c.km + 6) * (8 * c.nx * c.ny + 8) + 16
