"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs only where the reference tree is present: it drives oracle/_ref/qmiv_ref (the reference's
pmbb-core + pmbb-ivqm sources built for x86-64-v2 without FMA, linked with oracle/ref_harness.cpp)
on seeded synthetic inputs and on the reference's example frames, and stores inputs' SHA-256
plus every output as JSON. The example YUV files themselves are not copied; only the numbers
the reference computes on them are stored.

    python tests/golden/make_golden.py            # metric fixtures
    python tests/golden/make_golden.py --png      # PNG decoding (oracle/_ref/libspng_ref.so)
    python tests/golden/make_golden.py --multiview   # 24-view frame sets (4096x2048 ERP, 1920x1080, 2048x1088)
    python tests/golden/make_golden.py --ssimsets    # SSIM-family frame sets (16 x 1080p10, mixed sizes)
    python tests/golden/make_golden.py --formats     # 4:2:2, odd-width 4:2:0, 4:0:0 masks, out-of-range pels
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from ivpsnr_amd.synth import make_mask, make_pair, sha256  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref", "qmiv_ref")
EXAMPLES = "/root/reference/example"


def rows_hash(rows) -> str:
    return hashlib.sha256(np.asarray(rows, dtype="<u8").tobytes()).hexdigest()


def run_ref(w, h, bd, cf, f0, f1, metrics, erp=False, extra=()):
    cmd = [REF, "-w", str(w), "-h", str(h), "-bd", str(bd), "-cf", str(cf), "-i0", f0, "-i1", f1,
           "-ml", metrics, "-nth", "8"] + (["-erp", "360", "180"] if erp else []) + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True, check=True)
    # xPicP::check prints its findings on stdout (libs/pmbb-core/src/xPic.cpp:62-82), inside the
    # harness's JSON text: split them off, keeping them as the check's report.
    import re
    report = re.findall(r"(?:FILE BROKEN|DETECTED)[^\n]*\n", r.stdout)
    out = json.loads(re.sub(r"(?:FILE BROKEN|DETECTED)[^\n]*\n", "", r.stdout))
    if report:
        out["check_report"] = "".join(report)
    return out


def write_pairs(td, pairs):
    f0, f1 = os.path.join(td, "p0.yuv"), os.path.join(td, "p1.yuv")
    with open(f0, "wb") as a, open(f1, "wb") as b:
        for p in pairs:
            a.write(p.ref.tobytes())
            b.write(p.tst.tobytes())
    return f0, f1


def compact_frame(fr, keep_rows: bool):
    out = {}
    for k, v in fr.items():
        if k in ("cps_rows", "ws_rows"):
            arr = np.asarray(v, dtype=np.uint64)
            out[k + "_sha256"] = [[rows_hash(c) for c in d] for d in arr] if k == "cps_rows" else \
                [rows_hash(c) for c in arr]
            out[k + "_sum"] = arr.sum(axis=-1).tolist()
            if keep_rows:
                out[k] = arr.tolist()
        elif k in ("scp0", "scp1"):
            out[k + "_sha256"] = [hashlib.sha256(bytes.fromhex(x)).hexdigest() for x in v]
        elif k == "ws_weights":
            if keep_rows:
                out[k] = v
        else:
            out[k] = v
    return out


SMALL_CASES = [
    # w, h, bd, chroma, seed, generator knobs
    (64, 48, 8, 420, 1, {}),
    (130, 66, 10, 420, 2, {}),
    (67, 45, 8, 444, 3, {}),
    (256, 128, 10, 420, 4, {"luma_offset": 12.0}),
    (96, 70, 10, 420, 11, {}),
    (200, 40, 12, 420, 5, {}),
]


COLOUR_W, COLOUR_H = 56, 40


def colour_input(bd, seed):
    """Seeded planar uint16 planes (3, 40, 56) with samples in [0, 2^bd)."""
    import numpy as np
    return np.random.default_rng(seed).integers(0, 1 << bd, size=(3, COLOUR_H, COLOUR_W), dtype=np.uint16)


def colour_fixtures():
    """The reference's Q16 colour matrices and its conversions of seeded planes
    (xColorSpaceSTD::ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32)."""
    import hashlib
    import numpy as np
    tables = json.loads(subprocess.run([REF, "-cstables"], capture_output=True, text=True, check=True).stdout)
    cases = []
    with tempfile.TemporaryDirectory() as td:
        for bd in (8, 10, 14, 16):
            for space in range(4):
                for to_rgb in (0, 1):
                    seed = 1000 * bd + 10 * space + to_rgb
                    src = colour_input(bd, seed)
                    fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
                    src.astype("<u2").tofile(fi)
                    subprocess.run([REF, "-cs", str(to_rgb), str(space), str(COLOUR_W), str(COLOUR_H), str(bd), fi, fo],
                                   check=True)
                    out = np.fromfile(fo, "<u2")
                    cases.append({"bit_depth": bd, "space": space, "to_rgb": to_rgb, "seed": seed,
                                  "width": COLOUR_W, "height": COLOUR_H,
                                  "sha256_out": hashlib.sha256(out.tobytes()).hexdigest(),
                                  "head": out[:8].tolist()})
    return {"tables": tables, "cases": cases}


MASK_CASES = [
    # w, h, bd, chroma, seed, pair knobs, mask bit depth, mask chroma, mask kind
    (64, 48, 8, 420, 1, {}, 8, 420, "objects"),
    (130, 66, 10, 420, 2, {}, 8, 420, "objects"),
    (67, 45, 8, 444, 3, {}, 8, 444, "soft"),
    (256, 128, 10, 420, 4, {"luma_offset": 12.0}, 8, 420, "soft"),
    (96, 70, 16, 420, 5, {}, 16, 420, "soft"),
    (64, 48, 10, 420, 6, {}, 8, 420, "zero"),
    (64, 48, 10, 420, 7, {}, 10, 444, "full"),
    (1920, 1080, 8, 420, 8, {}, 8, 420, "objects"),
]


def masked_fixtures():
    """Masked mode (PSNR, WS-PSNR, IV-PSNR with a mask picture) by the reference's processors."""
    cases = []
    with tempfile.TemporaryDirectory() as td:
        for (w, h, bd, cf, seed, kw, bdm, cfm, kind) in MASK_CASES:
            pair = make_pair(w, h, bd, seed, chroma_format=cf, **kw)
            mask = make_mask(w, h, bdm, seed, chroma_format=cfm, kind=kind)
            f0, f1 = write_pairs(td, [pair])
            fm = os.path.join(td, "m.yuv")
            mask.tofile(fm)
            entry = {"width": w, "height": h, "bit_depth": bd, "chroma_format": cf, "seed": seed, "knobs": kw,
                     "mask_bit_depth": bdm, "mask_chroma_format": cfm, "mask_kind": kind,
                     "sha256_ref": sha256(pair.ref), "sha256_tst": sha256(pair.tst), "sha256_mask": sha256(mask)}
            for erp in (False, True):
                out = run_ref(w, h, bd, cf, f0, f1, "PSNR,WSPSNR,IVPSNR", erp,
                              ["-rows", "-im", fm, "-bdm", str(bdm), "-cfm", str(cfm)])
                entry["erp" if erp else "plain"] = compact_frame(out["frames"][0], keep_rows=h <= 128)
            cases.append(entry)
            print("masked", w, h, bd, cf, kind, file=sys.stderr)
    return cases


# Config 4: 24 views of uneven size, 10-bit 4:2:0; view i is the synthetic pair of seed 11 + i and
# its frame f that pair circularly shifted by 64 f samples (the benchmark's multi-view frame sets).
MULTIVIEW = [(4096, 2048, True)] * 8 + [(1920, 1080, False)] * 8 + [(2048, 1088, False)] * 8
MULTIVIEW_FRAMES = 2
MULTIVIEW_METRICS = "PSNR,WSPSNR,IVPSNR,IVSSIM"


def multiview_frame(pair, f):
    return np.roll(pair.ref, 64 * f), np.roll(pair.tst, 64 * f)


def multiview_fixtures():
    """The 24 views' first two frame sets (IV-PSNR, WS-PSNR, PSNR, IV-SSIM) by the reference."""
    views = []
    with tempfile.TemporaryDirectory() as td:
        for i, (w, h, erp) in enumerate(MULTIVIEW):
            pair = make_pair(w, h, 10, seed=11 + i)
            f0, f1 = os.path.join(td, "p0.yuv"), os.path.join(td, "p1.yuv")
            sh = []
            with open(f0, "wb") as a, open(f1, "wb") as b:
                for f in range(MULTIVIEW_FRAMES):
                    r, t = multiview_frame(pair, f)
                    a.write(r.tobytes())
                    b.write(t.tobytes())
                    sh.append([sha256(r), sha256(t)])
            out = run_ref(w, h, 10, 420, f0, f1, MULTIVIEW_METRICS, erp, ["-rows"])
            views.append({"width": w, "height": h, "bit_depth": 10, "erp": erp, "seed": 11 + i, "sha256": sh,
                          "frames": [compact_frame(fr, keep_rows=False) for fr in out["frames"]]})
            print("multiview", i, w, h, file=sys.stderr)
    return {"metrics": MULTIVIEW_METRICS, "views": views}


# SSIM-family frame sets: 16 frames 1920x1080 10-bit (seeds 101..116), and five frames of mixed
# sizes and bit depths (seeds 201..205), every metric.
SSIM_SET_1080 = [(1920, 1080, 10, 101 + f) for f in range(16)]
SSIM_SET_MIXED = [(1920, 1080, 10, 201), (640, 360, 8, 202), (2048, 1088, 10, 203), (1280, 720, 10, 204),
                  (352, 288, 8, 205)]


def ssim_set_fixtures():
    sets = {}
    with tempfile.TemporaryDirectory() as td:
        for name, frames, ml in (("1080p10", SSIM_SET_1080, "SSIM,MSSSIM,IVSSIM"), ("mixed", SSIM_SET_MIXED, "All")):
            out = []
            for (w, h, bd, seed) in frames:
                pair = make_pair(w, h, bd, seed=seed)
                f0, f1 = write_pairs(td, [pair])
                fr = run_ref(w, h, bd, 420, f0, f1, ml, False, ["-rows"])["frames"][0]
                fr = {k: v for k, v in compact_frame(fr, keep_rows=False).items() if k != "ssim_rows_cr"}
                out.append({"width": w, "height": h, "bit_depth": bd, "seed": seed, "sha256_ref": sha256(pair.ref),
                            "sha256_tst": sha256(pair.tst), "frame": fr})
                print("ssim set", name, w, h, bd, seed, file=sys.stderr)
            sets[name] = {"metrics": ml, "frames": out}
    return sets


def format_fixtures():
    """Raw-file formats the reference's reader unpacks into 4:4:4 (xSeq.cpp:179-249): 4:2:2, 4:2:0 of
    odd width, a 4:0:0 mask; and pictures holding pels above the bit depth, computed as read
    (InvalidPelActn WARN / SKIP) and concealed (CNCL). Inputs: tests/inputs.py."""
    from tests.inputs import FORMAT_CASES, MASK400_CASES, RANGE_CASES, format_pair, mask400_frame, range_mask, range_pair
    out = {"formats": [], "mask400": [], "range": []}
    with tempfile.TemporaryDirectory() as td:
        f0, f1, fm, fd = (os.path.join(td, n) for n in ("p0.yuv", "p1.yuv", "m.yuv", "dump.bin"))
        for (w, h, bd, fcf, seed) in FORMAT_CASES:
            a, b = format_pair(w, h, bd, fcf, seed)
            a.tofile(f0)
            b.tofile(f1)
            r = run_ref(w, h, bd, fcf, f0, f1, "All", False, ["-rows", "-dump", fd])
            planes = np.fromfile(fd, "<u2")
            out["formats"].append({"width": w, "height": h, "bit_depth": bd, "chroma_format": fcf, "seed": seed,
                                   "sha256_p0": hashlib.sha256(a.tobytes()).hexdigest(),
                                   "sha256_unpacked": hashlib.sha256(planes.tobytes()).hexdigest(),
                                   "frame": compact_frame(r["frames"][0], keep_rows=False)})
            print("format", w, h, bd, fcf, file=sys.stderr)
        for (w, h, bd, bdm, seed) in MASK400_CASES:
            pair = make_pair(w, h, bd, seed)
            write_pairs(td, [pair])
            mf = mask400_frame(w, h, bdm, seed)
            mf.tofile(fm)
            e = {"width": w, "height": h, "bit_depth": bd, "mask_bit_depth": bdm, "seed": seed,
                 "sha256_mask": hashlib.sha256(mf.tobytes()).hexdigest()}
            for erp in (False, True):
                r = run_ref(w, h, bd, 420, f0, f1, "PSNR,WSPSNR,IVPSNR", erp,
                            ["-rows", "-im", fm, "-bdm", str(bdm), "-cfm", "400"])
                e["erp" if erp else "plain"] = compact_frame(r["frames"][0], keep_rows=False)
            out["mask400"].append(e)
            print("mask400", w, h, bd, bdm, file=sys.stderr)
        for (w, h, bd, cf, seed, inj0, inj1, bdm) in RANGE_CASES:
            a, b = range_pair(w, h, bd, cf, seed, inj0, inj1)
            a.tofile(f0)
            b.tofile(f1)
            e = {"width": w, "height": h, "bit_depth": bd, "chroma_format": cf, "seed": seed, "inject": [inj0, inj1],
                 "mask_bit_depth": bdm, "sha256_p0": sha256(a), "sha256_p1": sha256(b)}
            extra = ["-rows"]
            ml = "All"
            if bdm:
                mk = range_mask(w, h, bdm, cf, seed)
                mk.tofile(fm)
                e["sha256_mask"] = sha256(mk)
                extra += ["-im", fm, "-bdm", str(bdm), "-cfm", str(cf)]
                ml = "PSNR,WSPSNR,IVPSNR"
            for mode in ("WARN", "CNCL"):
                r = run_ref(w, h, bd, cf, f0, f1, ml, False, extra + ["-ipa", mode])
                e[mode] = compact_frame(r["frames"][0], keep_rows=False)
            # The check's report with the file names as placeholders ({i0}, {i1}, {im}).
            rep = r.get("check_report", "").replace(f0, "{i0}").replace(f1, "{i1}").replace(fm, "{im}")
            e["check_report"] = rep
            out["range"].append(e)
            print("range", w, h, bd, cf, file=sys.stderr)
    return out


def png_fixtures():
    """The reference's PNG read step (libspng as xSeqPNG calls it, oracle/spng_ref.c) on the encoded
    cases of tests/png_writer.py: decode status and the SHA-256 of the R, G, B planes."""
    import ctypes as C
    from tests import png_writer
    spng = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libspng_ref.so"))
    out = []
    with tempfile.TemporaryDirectory() as td:
        for case in png_writer.cases():
            data, (w, h), _ = png_writer.build_case(case)
            path = os.path.join(td, case["name"] + ".png")
            with open(path, "wb") as f:
                f.write(data)
            planes = np.zeros(3 * w * h, np.uint8)
            err = C.create_string_buffer(256)
            rc = spng.spng_ref_decode(path.encode(), w, h, planes.ctypes.data_as(C.c_void_p), err, 256)
            out.append({"name": case["name"], "ok": rc == 0, "error": err.value.decode(),
                        "sha256": hashlib.sha256(planes.tobytes()).hexdigest() if rc == 0 else None})
    return {"cases": out}


def main():
    if "--png" in sys.argv:
        with open(os.path.join(HERE, "png.json"), "w") as f:
            json.dump(png_fixtures(), f, indent=0)
        return
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/qmiv_ref is not built (make -C oracle ref)")
    if "--masked" in sys.argv:
        with open(os.path.join(HERE, "masked.json"), "w") as f:
            json.dump(masked_fixtures(), f)
        return
    if "--multiview" in sys.argv:
        with open(os.path.join(HERE, "multiview24.json"), "w") as f:
            json.dump(multiview_fixtures(), f, separators=(",", ":"))
        return
    if "--formats" in sys.argv:
        with open(os.path.join(HERE, "formats.json"), "w") as f:
            json.dump(format_fixtures(), f, separators=(",", ":"))
        return
    if "--ssimsets" in sys.argv:
        with open(os.path.join(HERE, "ssim_sets.json"), "w") as f:
            json.dump(ssim_set_fixtures(), f, separators=(",", ":"))
        return
    if "--colour" in sys.argv:
        with open(os.path.join(HERE, "colorspace.json"), "w") as f:
            json.dump(colour_fixtures(), f)
        return
    fixtures = {}
    with tempfile.TemporaryDirectory() as td:
        # 1. Small synthetic cases: every metric, full per-row arrays, both WS settings.
        cases = []
        for (w, h, bd, cf, seed, kw) in SMALL_CASES:
            pair = make_pair(w, h, bd, seed, chroma_format=cf, **kw)
            f0, f1 = write_pairs(td, [pair])
            entry = {"width": w, "height": h, "bit_depth": bd, "chroma_format": cf, "seed": seed, "knobs": kw,
                     "sha256_ref": sha256(pair.ref), "sha256_tst": sha256(pair.tst)}
            for erp in (False, True):
                out = run_ref(w, h, bd, cf, f0, f1, "All", erp, ["-rows", "-pics"])
                entry["erp" if erp else "plain"] = compact_frame(out["frames"][0], keep_rows=True)
            cases.append(entry)
            print("small", w, h, bd, cf, file=sys.stderr)
        fixtures["small_cases"] = cases

        # 2. Plain IV-PSNR + PSNR: 3 frames 1920x1080 8-bit (seeds 1..3).
        pairs = [make_pair(1920, 1080, 8, seed=1 + f) for f in range(3)]
        f0, f1 = write_pairs(td, pairs)
        out = run_ref(1920, 1080, 8, 420, f0, f1, "PSNR,IVPSNR", False, ["-rows"])
        fixtures["plain_1080p8"] = {
            "width": 1920, "height": 1080, "bit_depth": 8, "chroma_format": 420, "seeds": [1, 2, 3],
            "sha256_ref": [sha256(p.ref) for p in pairs], "sha256_tst": [sha256(p.tst) for p in pairs],
            "frames": [compact_frame(fr, keep_rows=False) for fr in out["frames"]], "avg": out["avg"]}
        print("plain_1080p8", file=sys.stderr)

        # 3. IV-PSNR + WS-PSNR on ERP: one 4096x2048 10-bit frame (seed 1).
        pair = make_pair(4096, 2048, 10, seed=1)
        f0, f1 = write_pairs(td, [pair])
        out = run_ref(4096, 2048, 10, 420, f0, f1, "PSNR,WSPSNR,IVPSNR", True, ["-rows"])
        fixtures["erp_4k10"] = {
            "width": 4096, "height": 2048, "bit_depth": 10, "chroma_format": 420, "seed": 1,
            "sha256_ref": sha256(pair.ref), "sha256_tst": sha256(pair.tst),
            "frame": compact_frame(out["frames"][0], keep_rows=False)}
        print("erp_4k10", file=sys.stderr)

        # 4. SSIM family: one 1920x1080 10-bit frame (seed 1).
        pair = make_pair(1920, 1080, 10, seed=1)
        f0, f1 = write_pairs(td, [pair])
        out = run_ref(1920, 1080, 10, 420, f0, f1, "SSIM,MSSSIM,IVSSIM", False, ["-rows"])
        fixtures["ssim_1080p10"] = {
            "width": 1920, "height": 1080, "bit_depth": 10, "chroma_format": 420, "seed": 1,
            "sha256_ref": sha256(pair.ref), "sha256_tst": sha256(pair.tst),
            "frame": compact_frame(out["frames"][0], keep_rows=False)}
        print("ssim_1080p10", file=sys.stderr)

    # 5. The reference's example frames (1 frame 1920x1080 8-bit each); numbers only.
    if os.path.isdir(EXAMPLES):
        ex = {}
        orig = os.path.join(EXAMPLES, "Fencing_v8_original_1920x1080_yuv420p.yuv")
        for name in ("origJpgCompressed", "synthesized", "original"):
            other = os.path.join(EXAMPLES, f"Fencing_v8_{name}_1920x1080_yuv420p.yuv")
            out = run_ref(1920, 1080, 8, 420, orig, other, "All", False, ["-rows"])
            with open(orig, "rb") as a, open(other, "rb") as b:
                sh = [hashlib.sha256(a.read()).hexdigest(), hashlib.sha256(b.read()).hexdigest()]
            ex[name] = {"sha256_input0": sh[0], "sha256_input1": sh[1],
                        "frame": compact_frame(out["frames"][0], keep_rows=False)}
        # Published values (example/README.md:46-47), four decimals.
        ex["published"] = {"origJpgCompressed": {"psnr_y": 34.3362, "ivpsnr": 39.1335},
                           "synthesized": {"psnr_y": 34.0128, "ivpsnr": 44.3555}}
        fixtures["examples"] = ex
        print("examples", file=sys.stderr)

    # 6. The SSIM window as stored by the reference (fp32 bit patterns, 11 x 11).
    r = subprocess.run([REF, "-filter"], capture_output=True, text=True, check=True)
    fixtures["ssim_filter_bits"] = json.loads(r.stdout)

    for name, data in fixtures.items():
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(data, f, separators=(",", ":"))
    print("wrote", sorted(fixtures), file=sys.stderr)


if __name__ == "__main__":
    main()
