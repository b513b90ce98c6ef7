"""Malformed inputs to the file readers (host code, no device): every
truncation and a few hundred seeded byte corruptions of TIFF files
(uncompressed, PackBits, LZW with the predictor, both byte orders, several
pages and strips), of a MayaCam 2.0 calibration, a .vie filter file, a
trial .cfg and a .tra result file. Each read must either succeed or return
an error status: no crash, no abort, no runaway allocation. The reads run
in a child process so that a crash fails this test instead of the run. The
same corpus also goes through the readers built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpp/reader_fuzz.cpp).
"""
import os
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLD = Path(__file__).resolve().parent / "golden" / "sample_data"


def test_readers_survive_malformed_files(tmp_path):
    script = tmp_path / "fuzz.py"
    script.write_text(textwrap.dedent(f"""
        import sys, random
        from pathlib import Path
        sys.path[:0] = [{str(ROOT)!r}, {str(ROOT / 'tests')!r}]
        import numpy as np
        import autoscoper_amd as A
        from test_io import write_tiff

        tmp = Path({str(tmp_path)!r})
        rng = random.Random(20261203)
        counts = {{"ok": 0, "error": 0}}

        def attempt(fn):
            try:
                fn()
                counts["ok"] += 1
            except A.AscError:
                counts["error"] += 1

        def read_tiff(p):
            w, h, n, b = A.io.tiff_info(p)
            if w * h * n * b // 8 <= (16 << 20):
                A.io.read_tiff(p)

        variants = []
        a8 = np.arange(7 * 9, dtype=np.uint8).reshape(7, 9)
        a16 = (np.arange(6 * 10, dtype=np.uint16) * 997).reshape(6, 10)
        for i, (pages, kw) in enumerate([
                ([a8, a8 + 1, a8 + 2], dict()),
                ([a16], dict(compression=32773, order=">", rows_per_strip=2)),
                ([a16, a16], dict(compression=5, predictor=2, rows_per_strip=4)),
                ([a8], dict(compression=5))]):
            p = tmp / f"v{{i}}.tif"
            write_tiff(p, pages, **kw)
            variants.append(p.read_bytes())
        bad = tmp / "bad.tif"
        for b in variants:
            for cut in range(len(b)):
                bad.write_bytes(b[:cut])
                attempt(lambda: read_tiff(bad))
            for _ in range(300):
                c = bytearray(b)
                for _ in range(rng.randint(1, 4)):
                    c[rng.randrange(len(c))] = rng.randrange(256)
                bad.write_bytes(bytes(c))
                attempt(lambda: read_tiff(bad))

        texts = [({str(GOLD / 'Calibration' / 'xr_calib_wrist_cam01.txt')!r}, "cam.txt", A.io.read_camera),
                 ({str(GOLD / 'xParameters' / 'control_settings.vie')!r}, "f.vie", A.io.read_filter_settings)]
        for src, name, fn in texts:
            b = Path(src).read_bytes()
            p = tmp / name
            for cut in range(0, len(b), max(1, len(b) // 200)):
                p.write_bytes(b[:cut])
                attempt(lambda: fn(p))
            for _ in range(300):
                c = bytearray(b)
                for _ in range(rng.randint(1, 4)):
                    c[rng.randrange(len(c))] = rng.choice(b"0123456789.-+eE ,\\n\\tx")
                p.write_bytes(bytes(c))
                attempt(lambda: fn(p))

        # the trial's relative paths resolve beside a copy of its .cfg
        gold = Path({str(GOLD)!r})
        for d in ("Calibration", "Models", "XMA_UND", "xParameters"):
            (tmp / d).symlink_to(gold / d)
        cfg = (gold / "wrist.cfg").read_text()
        p = tmp / "fuzz.cfg"
        for _ in range(200):
            lines = cfg.splitlines()
            k = rng.randrange(len(lines))
            words = lines[k].split()
            if words:
                words[rng.randrange(len(words))] = rng.choice(["", "-1", "1e400", "nan", "../x", "0 0 0"])
            lines[k] = " ".join(words)
            p.write_text("\\n".join(lines))
            attempt(lambda: A.Trial(p).close())

        good = np.arange(3 * 2 * 6, dtype=float).reshape(3, 2, 6)
        tra = tmp / "r.tra"
        A.io.write_tra(tra, good)
        b = tra.read_bytes()
        for _ in range(200):
            c = bytearray(b)
            for _ in range(rng.randint(1, 3)):
                c[rng.randrange(len(c))] = rng.choice(b"0123456789.-,e\\nNa ")
            tra.write_bytes(bytes(c))
            attempt(lambda: A.io.read_tra(tra, 3, 2))
        print("fuzz", counts, flush=True)
    """))
    out = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, (out.returncode, out.stdout[-2000:], out.stderr[-4000:])
    assert "fuzz" in out.stdout
    print(out.stdout.strip())


def _corpus(tmp, asc):
    """Corrupted copies of every input kind, written under tmp/<kind>/."""
    import random

    from test_io import write_tiff

    rng = random.Random(7)
    files = {k: [] for k in ("tiff", "cam", "vie", "tra", "cfg")}
    for k in files:
        (tmp / k).mkdir()
    a8 = np.arange(7 * 9, dtype=np.uint8).reshape(7, 9)
    a16 = (np.arange(6 * 10, dtype=np.uint16) * 997).reshape(6, 10)
    for i, (pages, kw) in enumerate([([a8, a8 + 1, a8 + 2], {}),
                                     ([a16], dict(compression=32773, order=">", rows_per_strip=2)),
                                     ([a16, a16], dict(compression=5, predictor=2, rows_per_strip=4)),
                                     ([a8], dict(compression=5))]):
        base = tmp / f"base{i}.tif"
        write_tiff(base, pages, **kw)
        b = base.read_bytes()
        variants = [b[:cut] for cut in range(len(b))]
        for _ in range(500):
            c = bytearray(b)
            for _ in range(rng.randint(1, 6)):
                c[rng.randrange(len(c))] = rng.randrange(256)
            variants.append(bytes(c))
        for j, v in enumerate(variants):
            p = tmp / "tiff" / f"{i}_{j}.tif"
            p.write_bytes(v)
            files["tiff"].append(p)
    for src, kind in ((GOLD / "Calibration" / "xr_calib_wrist_cam01.txt", "cam"),
                      (GOLD / "xParameters" / "control_settings.vie", "vie")):
        b = src.read_bytes()
        for j in range(500):
            c = bytearray(b[: rng.randrange(len(b))] if j < 100 else b)
            for _ in range(rng.randint(1, 6) if j >= 100 else 0):
                c[rng.randrange(len(c))] = rng.choice(b"0123456789.-+eE ,\n\tx")
            p = tmp / kind / f"{j}.txt"
            p.write_bytes(bytes(c))
            files[kind].append(p)
    base = tmp / "base.tra"
    asc.io.write_tra(base, np.arange(3 * 2 * 6, dtype=float).reshape(3, 2, 6))
    b = base.read_bytes()
    for j in range(500):
        c = bytearray(b)
        for _ in range(rng.randint(1, 4)):
            c[rng.randrange(len(c))] = rng.choice(b"0123456789.-,e\nNa ")
        p = tmp / "tra" / f"{j}.tra"
        p.write_bytes(bytes(c))
        files["tra"].append(p)
    for d in ("Calibration", "Models", "XMA_UND", "xParameters"):
        (tmp / "cfg" / d).symlink_to(GOLD / d)
    cfg = (GOLD / "wrist.cfg").read_text()
    for j in range(200):
        lines = cfg.splitlines()
        k = rng.randrange(len(lines))
        words = lines[k].split()
        if words:
            words[rng.randrange(len(words))] = rng.choice(["", "-1", "1e400", "nan", "../x", "0 0 0", "99999999999"])
        lines[k] = " ".join(words)
        p = tmp / "cfg" / f"{j}.cfg"
        p.write_text("\n".join(lines))
        files["cfg"].append(p)
    return files


def test_readers_under_sanitizers(asc, tmp_path):
    """The host readers (io.cpp), the remote session (remote.cpp) and the
    pose math they use (geometry.cpp), built with -fsanitize=address,undefined
    into tests/cpp/reader_fuzz.cpp, replay the corrupted corpus and random
    protocol messages: any out-of-bounds access, use of freed memory or
    undefined behaviour aborts with the sanitizer's report."""
    import shutil

    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    src = ROOT / "autoscoper_amd" / "csrc"
    exe = tmp_path / "reader_fuzz"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-ffp-contract=off", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "reader_fuzz.cpp"),
           str(src / "io.cpp"), str(src / "remote.cpp"), str(src / "geometry.cpp"),
           f"-L{ROOT / 'autoscoper_amd'}", "-lautoscoper_hip", f"-Wl,-rpath,{ROOT / 'autoscoper_amd'}", "-lpthread",
           "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if b.returncode != 0 and "asan" in b.stderr.lower():
        pytest.skip("the sanitizer runtime is not installed: " + b.stderr[-300:])
    assert b.returncode == 0, b.stderr[-3000:]
    files = _corpus(tmp_path, asc)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    for kind, paths in files.items():
        out = subprocess.run([str(exe), kind, *map(str, paths)], capture_output=True, text=True, timeout=600,
                             env=env, cwd=tmp_path)
        assert out.returncode == 0 and "runtime error" not in out.stderr, (kind, out.stderr[-4000:])
        print(kind, out.stdout.strip())
    work = tmp_path / "session"
    work.mkdir()
    out = subprocess.run([str(exe), "messages", str(GOLD / "wrist.cfg"), "10000"], capture_output=True, text=True,
                         timeout=600, env=env, cwd=work)
    assert out.returncode == 0 and "runtime error" not in out.stderr, out.stderr[-4000:]
    print("messages", out.stdout.strip())
