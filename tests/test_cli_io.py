"""Tests of the drop-in surface: OBJ / glTF / GLB loading and the RT-RCS
command line up to the device step (config keys, CLI rules, error exits) on
the CPU, and one RT-RCS sweep end to end on the GPU."""
import base64
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import sagittasbr_amd as S
from sagittasbr_amd import meshes as M

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
RTRCS = os.path.join(ROOT, "sagittasbr_amd", "RT-RCS")


def test_obj_number_grammar():
    P = S.obj_parse_number
    assert P("11e2") == 1100.0 and P("-0") == 0.0 and P("42") == 42.0
    assert P("0.5") == 0.5 and P(".7e+2") == 70.0 and P("-.5") == -0.5
    assert P("+3.1417e+2") == pytest.approx(314.17, rel=1e-15)
    assert P("1e") is None and P("abc") is None and P("") is None
    assert P("2.5x") == 2.5  # greedy: stops at the first non-conforming char


def test_obj_numbers_close_to_correct_rounding():
    """tinyobj's accumulation is not correctly rounded, but once narrowed to
    float it stays within one float ulp of the correctly rounded value."""
    rng = np.random.default_rng(1)
    vals = rng.uniform(-50, 50, 2000).astype(np.float32)
    same = 0
    for x in vals:
        tok = f"{float(x):.9g}"
        got = np.float32(S.obj_parse_number(tok))
        ref = np.float32(float(tok))
        assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1
        same += got == ref
    assert same > 0.9 * len(vals)


def test_obj_box_and_sphere_fixtures():
    v, f = S.load_model(os.path.join(HERE, "golden", "ref_box.obj"))
    assert f.shape == (12, 3) and v.shape == (8, 3)
    assert set(np.unique(v).tolist()) == {-1.0, 1.0}
    v2, f2 = S.load_model(os.path.join(HERE, "golden", "ref_sphere.obj"))
    assert f2.shape[0] == 1280
    r = np.linalg.norm(v2, axis=1)
    assert np.all(np.abs(r - 1) < 1e-5)
    # positions equal those of a straightforward reader up to one float ulp
    rv, rf = M.read_obj(os.path.join(HERE, "golden", "ref_sphere.obj"))
    np.testing.assert_array_equal(rv[rf.astype(np.int64)].shape, v2[f2.astype(np.int64)].shape)
    assert np.max(np.abs(rv[rf.astype(np.int64)] - v2[f2.astype(np.int64)])) <= 1.2e-7


def test_obj_quads_split_on_shorter_diagonal_and_negative_indices(tmp_path):
    p = tmp_path / "q.obj"
    # quad 0: diagonal 0-2 shorter -> (0,1,2),(0,2,3); quad 1: 1-3 shorter -> (0,1,3),(1,2,3)
    p.write_text("v 0 0 0\nv 2 0 0\nv 1.5 1 0\nv -0.5 1 0\n"
                 "v 10 0 0\nv 12 0 0\nv 12 1 0\nv 10.3 1.2 0\n"
                 "f 1 2 3 4\nf -4 -3 -2 -1\nf 1/1/1 2/2/2 3/3/3\n")
    v, f = S.load_model(str(p))
    tris = v[f.astype(np.int64)]
    want = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 7], [5, 6, 7], [0, 1, 2]])
    vv = np.array([[0, 0, 0], [2, 0, 0], [1.5, 1, 0], [-0.5, 1, 0], [10, 0, 0], [12, 0, 0], [12, 1, 0], [10.3, 1.2, 0]],
                  np.float32)
    np.testing.assert_array_equal(tris, vv[want])


def _gltf_doc(verts, faces, uri=None, pos_type=5126, mode=4):
    vb = np.asarray(verts, np.float32).tobytes()
    ib = np.asarray(faces, np.uint16).reshape(-1).tobytes()
    pad = (4 - len(ib) % 4) % 4
    blob = ib + b"\0" * pad + vb
    doc = {
        "asset": {"version": "2.0"},
        "buffers": [{"byteLength": len(blob)}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": len(ib)},
                        {"buffer": 0, "byteOffset": len(ib) + pad, "byteLength": len(vb)}],
        "accessors": [{"bufferView": 0, "componentType": 5123, "count": int(np.size(faces)), "type": "SCALAR"},
                      {"bufferView": 1, "componentType": pos_type, "count": len(verts), "type": "VEC3"}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 1}, "indices": 0, "mode": mode}]}],
    }
    if uri is not None:
        doc["buffers"][0]["uri"] = uri(blob)
    return doc, blob


def test_gltf_data_uri_and_glb(tmp_path):
    v, f = M.icosphere(1)
    doc, blob = _gltf_doc(v, f, uri=lambda b: "data:application/octet-stream;base64," + base64.b64encode(b).decode())
    g = tmp_path / "s.gltf"
    g.write_text(json.dumps(doc))
    gv, gf = S.load_model(str(g))
    np.testing.assert_array_equal(gv, v)
    np.testing.assert_array_equal(gf, f)
    # binary container
    doc2, blob2 = _gltf_doc(v, f)
    js = json.dumps(doc2).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    bb = blob2 + b"\0" * ((4 - len(blob2) % 4) % 4)
    glb = struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(bb))
    glb += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bb), 0x004E4942) + bb
    p = tmp_path / "s.glb"
    p.write_bytes(glb)
    bv, bf = S.load_model(str(p))
    np.testing.assert_array_equal(bv, v)
    np.testing.assert_array_equal(bf, f)
    # external .bin buffer
    (tmp_path / "s.bin").write_bytes(blob)
    doc3, _ = _gltf_doc(v, f, uri=lambda b: "s.bin")
    (tmp_path / "e.gltf").write_text(json.dumps(doc3))
    ev, ef = S.load_model(str(tmp_path / "e.gltf"))
    np.testing.assert_array_equal(ev, v)


def test_gltf_skips_non_float_positions_and_non_triangles(tmp_path):
    v, f = M.icosphere(0)
    for kw in (dict(pos_type=5123), dict(mode=1)):
        doc, blob = _gltf_doc(v, f, uri=lambda b: "data:;base64," + base64.b64encode(b).decode(), **kw)
        g = tmp_path / "x.gltf"
        g.write_text(json.dumps(doc))
        gv, gf = S.load_model(str(g))
        assert len(gf) == 0


def _run(args, cwd):
    return subprocess.run([RTRCS] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(not os.path.exists(RTRCS), reason="RT-RCS not built")
def test_cli_config_and_flags(tmp_path):
    (tmp_path / "c.txt").write_text("# comment\nphi_start 5 # trailing\nphi_samples 7\ntheta_end 45\nn_x 123\n"
                                    "freq 2e9\nmax_bounces 3\nunknown_key 4\n")
    r = _run(["--config", "c.txt", "--print-config", "--model", "m.obj"], tmp_path)
    assert r.returncode == 0
    assert "Using frequency from config file: 2e+09 Hz" in r.stdout
    kv = dict(x.split("=", 1) for x in r.stdout.strip().splitlines()[-1].split())
    assert kv["phi_start"] == "5" and kv["phi_samples"] == "7" and kv["theta_end"] == "45"
    assert kv["n_x"] == "123" and kv["n_y"] == "5000" and kv["max_bounces"] == "3"
    r = _run(["--config", "c.txt", "--print-config", "3.5e9", "4e9", "--bogus"], tmp_path)
    assert "Using frequency from command line: 3.5e+09 Hz" in r.stdout
    assert "Warning: Unrecognized argument '--bogus'." in r.stderr
    r = _run(["--print-config"], tmp_path)
    assert "Warning: Could not open config.txt. Using defaults." in r.stderr
    assert "Using default frequency: 1e+10 Hz" in r.stdout


@pytest.mark.skipif(not os.path.exists(RTRCS), reason="RT-RCS not built")
def test_cli_error_exits(tmp_path):
    r = _run([], tmp_path)
    assert r.returncode == 1 and "Error: --model is required for simulation runs." in r.stderr
    r = _run(["--model", "x.stl"], tmp_path)
    assert r.returncode == 1 and "Unsupported model extension for 'x.stl'" in r.stderr
    r = _run(["--model", "missing.obj"], tmp_path)
    assert r.returncode == 1 and "Failed to load model 'missing.obj'" in r.stderr
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n")
    r = _run(["--model", "bad.obj"], tmp_path)
    assert r.returncode == 1
    r = _run(["--model", "--x"], tmp_path)  # --model consumes the next token
    assert r.returncode == 1
    r = _run(["--model"], tmp_path)
    assert r.returncode == 1 and "--model requires a file path" in r.stderr


def _read_csv(path):
    lines = open(path).read().splitlines()
    head = [x for x in lines if x.startswith("#")]
    assert lines[len(head)] == "theta,phi,rcs_m2,rcs_dbsm"
    rows = np.array([[float(t) for t in x.split(",")] for x in lines[len(head) + 1:]])
    return head, rows


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(RTRCS), reason="RT-RCS not built")
def test_cli_sweep_csv_matches_oracle(tmp_path):
    """RT-RCS end to end on the GPU: an OBJ model and a config.txt in, the
    rcs_results.csv rows out (rtRcsEntry.cu:298-302, sweep.cu:540-544),
    checked against the FP64 oracle sweep of the same loaded vertices. The
    CSV holds 6 significant digits, so sigma is compared to 6e-6 relative."""
    import oracle_lib as O

    v, f = M.trihedral_duct_scene()
    M.write_obj(str(tmp_path / "scene.obj"), v, f)
    lv, lf = S.load_model(str(tmp_path / "scene.obj"))
    kw = dict(theta_start=0, theta_end=90, theta_samples=4, phi_start=0, phi_end=90, phi_samples=5, grid_size=1.6,
              n_x=160, n_y=160, max_bounces=10, freq=10e9)
    (tmp_path / "config.txt").write_text("".join(f"{k} {kw[k]}\n" for k in kw if k != "freq") + "freq 1e10\n")
    n = kw["theta_samples"] * kw["phi_samples"]
    th = np.repeat(np.linspace(0, 90, 4), 5)
    ph = np.tile(np.linspace(0, 90, 5), 4)

    def sigma_db(z):
        s = 4 * np.pi * np.abs(z) ** 2
        return s, 10 * np.log10(np.maximum(s, 1e-10))

    for prec, method, scale in (("fp64", "batched", 1.0), ("fp64", "per_angle", 2.0), ("mixed", "batched", 1.0),
                                ("fp32", "batched", 1.0)):
        print(f"[cli] precision {prec} method {method} scale {scale}", flush=True)
        if prec == "fp32":
            # the oracle's FP32 twin traces every ray as the FP32 kernel does;
            # only the order of the FP32 sums differs
            z, _ = O.Scene(lv, lf, 0, scale=scale).sweep(O.Params(**kw), fused=True)
        else:
            z, _ = O.Scene(lv, lf, 1, scale=scale).sweep(O.Params(**kw))
        s_ref, db_ref = sigma_db(z)
        r = _run(["--model", "scene.obj", "--model-scale", str(scale), "--precision", prec, "--method", method],
                 tmp_path)
        assert r.returncode == 0, r.stderr[-2000:]
        head, rows = _read_csv(tmp_path / "rcs_results.csv")
        assert head == ["# Frequency: 1e+10", "# Grid: 160x160", "# GridSize: 1.6", "# ThetaSamples: 4",
                        "# PhiSamples: 5"]
        assert rows.shape == (n, 4)
        np.testing.assert_allclose(rows[:, 0], th, atol=1e-4)
        np.testing.assert_allclose(rows[:, 1], ph, atol=1e-4)
        if prec == "fp32":
            # |sqrt(sigma) - sqrt(sigma_ref)| <= sqrt(4 pi) |z - z_ref|: the
            # amplitudes agree to the rounding of two FP32 sums
            amp, amp_ref = np.sqrt(rows[:, 2]), np.sqrt(s_ref)
            assert np.max(np.abs(amp - amp_ref)) <= 1e-4 * amp_ref.max()
            continue
        np.testing.assert_allclose(rows[:, 2], s_ref, rtol=6e-6, atol=1e-12 * s_ref.max())
        np.testing.assert_allclose(rows[:, 3], db_ref, atol=2e-4)
