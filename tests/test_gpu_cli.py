"""The command-line driver on the GPU: every frame line, average line and result-file line it prints
equals the reference application's format (tests/cli_format.py) filled with the oracle's values for
the same frames. PSNR-family lines must match as text (bit-exact values); SSIM-family values are
parsed and compared within 1e-6. Covers start frames, ERP, the metric list, masked mode, YCbCr ->
RGB and BGR -> RGB conversions, identical components and PNG-list input."""
import numpy as np
import pytest

from ivpsnr_amd.synth import make_mask, make_pair
from tests.cli_format import avg_lines, frame_cmp_line, frame_pic_line, pic_value, run_cli, write_seq

pytestmark = pytest.mark.gpu

ALL = 63


def frames_for(w, h, bd, cf, n, seed, **knobs):
    return [make_pair(w, h, bd, seed + i, chroma_format=cf, **knobs) for i in range(n)]


def expected_psnr_family(results, names, mask=False, rgb=False, exact=None):
    """Frame and average lines of PSNR, WSPSNR and IVPSNR from per-frame oracle results."""
    from oracle import pyoracle
    lines, avg = [], []
    for name in names:
        cmp_rows, pics = [], []
        for f, r in enumerate(results):
            ex = exact[f] if exact is not None else list(r.exact)
            if name == "IVPSNR":
                pics.append(r.ivpsnr)
                lines.append(frame_pic_line(f, name, r.ivpsnr, mask, rgb))
                continue
            v = list(r.psnr if name == "PSNR" else r.wspsnr)
            cmp_rows.append(v)
            pics.append(pic_value(v))
            lines.append(frame_cmp_line(f, name, v, mask, rgb, ex))
            lines.append(frame_pic_line(f, name, pics[-1], mask, rgb))
        n = len(results)
        cmp_avg = [pyoracle.frame_average([c[i] for c in cmp_rows]) for i in range(3)] if cmp_rows else None
        avg.append((name, cmp_avg, pyoracle.frame_average(pics)))
    return lines, avg


def check_output(r, lines, avg, mask=False, rgb=False, result_file=None):
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    out = r.stdout.splitlines()
    for line in lines:
        assert line in out, line
    for name, cmp_avg, pic_avg in avg:
        for line in avg_lines("Average      ", name, cmp_avg, pic_avg, mask, rgb):
            assert line in out, line
    assert out[-1] == "END-OF-LOG"
    if result_file is not None:
        text = result_file.read_text().splitlines()
        assert text[0].startswith("FILE0  \"") and text[1].startswith("FILE1  \"")
        body = [l for l in text if not l.startswith(("FILE", "TIME"))]
        want = [l for name, c, p in avg for l in avg_lines("", name, c, p, mask, rgb)]
        assert body == want


def test_plain_all_metrics_start_frames_erp(tmp_path, oracle):
    w, h, bd = 128, 64, 10
    pairs = frames_for(w, h, bd, 420, 4, 100)
    write_seq(tmp_path / "ref.yuv", [p.ref for p in pairs], bd)
    write_seq(tmp_path / "tst.yuv", [p.tst for p in pairs[1:]], bd)  # tst frame k pairs with ref frame k + 1
    print("cli plain 128x64 bd10 erp all metrics", flush=True)
    r = run_cli(["-i0", "ref.yuv", "-i1", "tst.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-s0", 1, "-erp", "-ml", "All",
                 "-v", 2, "-r", "res.txt"], tmp_path)
    prm = oracle.params(w, h, bd, 420, metrics=ALL, erp=True)
    res = [oracle.frame(pairs[f + 1].ref, pairs[f + 1].tst, prm) for f in range(3)]
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR", "IVPSNR"])
    check_output(r, lines, avg)
    out = r.stdout.splitlines()
    frame_tokens = [l.split() for l in out if l.startswith("Frame ")]
    # normalized metrics: printed values within 1e-6 of the oracle's (6 decimals, SSIM within 1e-6)
    for name, attr in (("SSIM", "ssim"), ("MSSSIM", "msssim")):
        cmp_got = [list(map(float, t[4:7])) for t in frame_tokens if t[2] == name and t[3] == "Y:Cb:Cr"]
        pic_got = [float(t[3]) for t in frame_tokens if t[2] == name + "-YCbCr"]
        np.testing.assert_allclose(cmp_got, [list(getattr(x, attr)) for x in res], atol=1.5e-6)
        np.testing.assert_allclose(pic_got, [pic_value(list(getattr(x, attr))) for x in res], atol=1.5e-6)
    ivssim_got = [float(t[3]) for t in frame_tokens if t[2] == "IVSSIM"]
    np.testing.assert_allclose(ivssim_got, [x.ivssim for x in res], atol=1.5e-6)
    assert "FramesToProcess  = 3" in out
    assert len((tmp_path / "res.txt").read_text().splitlines()) == 3 + 2 + 2 + 1 + 2 + 2 + 1


def test_result_file_and_verbose0(tmp_path, oracle):
    w, h, bd = 96, 48, 8
    pairs = frames_for(w, h, bd, 444, 2, 200)
    write_seq(tmp_path / "a.yuv", [p.ref for p in pairs], bd)
    write_seq(tmp_path / "b.yuv", [p.tst for p in pairs], bd)
    print("cli 444 bd8 result file", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-pf", "yuv444p", "-ps", f"{w}x{h}", "-ml", "PSNR,WSPSNR,IVPSNR",
                 "-sr", 3, "-cws", "2:1:1:0", "-cwa", "4:1:1:0", "-v", 0, "-r", "res.txt"], tmp_path)
    prm = oracle.params(w, h, bd, 444, search_range=3, weights_search=(2, 1, 1, 0))
    res = [oracle.frame(p.ref, p.tst, prm) for p in pairs]
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR", "IVPSNR"])
    check_output(r, [], avg, result_file=tmp_path / "res.txt")
    assert not any(l.startswith("Frame ") for l in r.stdout.splitlines())


def test_masked(tmp_path, oracle):
    w, h, bd, bdm = 160, 96, 10, 8
    pairs = frames_for(w, h, bd, 420, 3, 300)
    masks = [make_mask(w, h, bdm, 300 + i, kind="soft" if i % 2 else "objects") for i in range(3)]
    write_seq(tmp_path / "a.yuv", [p.ref for p in pairs], bd)
    write_seq(tmp_path / "b.yuv", [p.tst for p in pairs], bd)
    write_seq(tmp_path / "m.yuv", masks, bdm)
    print("cli masked 160x96 bd10 mask bd8", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-im", "m.yuv", "-bdm", bdm,
                 "-ml", "PSNR, WSPSNR, IVPSNR", "-erp", "-v", 4, "-r", "res.txt"], tmp_path)
    prm = oracle.params(w, h, bd, 420, erp=True)
    res = [oracle.frame_masked(p.ref, p.tst, m, bdm, 420, prm) for p, m in zip(pairs, masks)]
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR"], mask=True)
    for f, x in enumerate(res):
        lines.append(frame_pic_line(f, "IVPSNR", x.ivpsnr, mask=True) + f"    R2T {x.r2t:7.4f}  T2R {x.t2r:7.4f}")
    avg.append(("IVPSNR", None, oracle.frame_average([x.ivpsnr for x in res])))
    check_output(r, lines, avg, mask=True, result_file=tmp_path / "res.txt")
    debug = [l for l in r.stdout.splitlines() if l.startswith("Frame") and "NNM" in l]
    assert [int(l.split()[3]) for l in debug] == [x.num_non_masked for x in res]
    assert [list(map(int, l.split()[5:9])) for l in debug] == [list(x.gcd) for x in res]


def pack444(planes, bd):
    return np.ascontiguousarray(planes.reshape(-1), np.uint8 if bd <= 8 else np.uint16)


@pytest.mark.parametrize("space_name,space", [("YCbCr_BT709", 1), ("YCbCr_BT2020", 3), ("YCbCr_SMPTE170M", 0)])
def test_ycbcr_to_rgb(tmp_path, oracle, space_name, space):
    w, h, bd = 128, 64, 10
    pairs = frames_for(w, h, bd, 420, 2, 400 + space)
    write_seq(tmp_path / "a.yuv", [p.ref for p in pairs], bd)
    write_seq(tmp_path / "b.yuv", [p.tst for p in pairs], bd)
    print(f"cli {space_name} -> RGB", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-csi", space_name, "-csm", "RGB",
                 "-ml", "PSNR,WSPSNR,IVPSNR", "-v", 2], tmp_path)
    prm = oracle.params(w, h, bd, 444)
    res = []
    for p in pairs:
        conv = [pack444(oracle.colorspace(oracle.unpack(x, w, h, bd, 420), bd, space, True), bd) for x in (p.ref, p.tst)]
        res.append(oracle.frame(conv[0], conv[1], prm))
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR", "IVPSNR"], rgb=True)
    check_output(r, lines, avg, rgb=True)


def test_bgr_reorder_and_rgb_to_ycbcr(tmp_path, oracle):
    w, h, bd = 64, 48, 12
    pairs = frames_for(w, h, bd, 444, 2, 500)
    write_seq(tmp_path / "a.yuv", [p.ref for p in pairs], bd)
    write_seq(tmp_path / "b.yuv", [p.tst for p in pairs], bd)
    prm = oracle.params(w, h, bd, 444)
    print("cli GBR -> RGB", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-cf", 444, "-csi", "GBR", "-csm",
                 "RGB", "-ml", "PSNR,IVPSNR", "-v", 2], tmp_path)
    gbr_to_rgb = lambda x: pack444(oracle.unpack(x, w, h, bd, 444)[[2, 0, 1]], bd)  # G,B,R planes -> R,G,B
    res = [oracle.frame(gbr_to_rgb(p.ref), gbr_to_rgb(p.tst), prm) for p in pairs]
    lines, avg = expected_psnr_family(res, ["PSNR", "IVPSNR"], rgb=True)
    check_output(r, lines, avg, rgb=True)
    print("cli RGB -> YCbCr_BT709", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-cf", 444, "-csi", "RGB", "-csm",
                 "YCbCr_BT709", "-ml", "PSNR,IVPSNR", "-v", 2], tmp_path)
    to_ycc = lambda x: pack444(oracle.colorspace(oracle.unpack(x, w, h, bd, 444), bd, 1, False), bd)
    res = [oracle.frame(to_ycc(p.ref), to_ycc(p.tst), prm) for p in pairs]
    lines, avg = expected_psnr_family(res, ["PSNR", "IVPSNR"])
    check_output(r, lines, avg)


def test_identical_components(tmp_path, oracle):
    """Identical chroma gets the PSNR of SSD=1 and the ExactU / ExactV tags (xAppQMIV.cpp:837-871). With
    a YCbCr -> RGB conversion the flags are those of the input pictures: identical luma alone marks R
    although R differs after conversion."""
    w, h, bd = 64, 32, 10
    n = w * h
    pair = make_pair(w, h, bd, 600)
    tst_chroma_same = pair.tst.copy()
    tst_chroma_same[n:] = pair.ref[n:]
    tst_luma_same = pair.tst.copy()
    tst_luma_same[:n] = pair.ref[:n]
    write_seq(tmp_path / "a.yuv", [pair.ref, pair.ref], bd)
    write_seq(tmp_path / "b.yuv", [tst_chroma_same, tst_luma_same], bd)
    fake = oracle.psnr_from_ssd(1.0, n, bd)
    print("cli identical components", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-ml", "PSNR,WSPSNR", "-v", 2], tmp_path)
    prm = oracle.params(w, h, bd, 420)
    res = [oracle.frame(pair.ref, t, prm) for t in (tst_chroma_same, tst_luma_same)]
    assert list(res[0].exact) == [0, 1, 1] and res[0].psnr[1] == fake
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR"])
    check_output(r, lines, avg)
    assert "ExactU ExactV" in r.stdout and "ExactY" in r.stdout

    print("cli identical luma with YCbCr_BT709 -> RGB", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-csi", "YCbCr_BT709", "-csm", "RGB",
                 "-ml", "PSNR", "-v", 2], tmp_path)
    prm444 = oracle.params(w, h, bd, 444)
    rows, exacts, pics = [], [], []
    lines = []
    for f, t in enumerate((tst_chroma_same, tst_luma_same)):
        conv = [pack444(oracle.colorspace(oracle.unpack(x, w, h, bd, 420), bd, 1, True), bd) for x in (pair.ref, t)]
        x = oracle.frame(conv[0], conv[1], prm444)
        ex = list(res[f].exact)  # flags of the input pictures
        v = [fake if ex[c] else x.psnr[c] for c in range(3)]
        rows.append(v)
        pics.append(pic_value(v))
        lines += [frame_cmp_line(f, "PSNR", v, rgb=True, exact=ex), frame_pic_line(f, "PSNR", pics[-1], rgb=True)]
    avg = [("PSNR", [oracle.frame_average([v[c] for v in rows]) for c in range(3)], oracle.frame_average(pics))]
    check_output(r, lines, avg, rgb=True)


def test_invalid_pel_stops(tmp_path):
    w, h, bd = 32, 16, 10
    pair = make_pair(w, h, bd, 700)
    bad = pair.tst.copy()
    bad[5] = 1500
    write_seq(tmp_path / "a.yuv", [pair.ref], bd)
    write_seq(tmp_path / "b.yuv", [bad], bd)
    print("cli invalid pel", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-ml", "PSNR"], tmp_path)
    assert r.returncode != 0 and "InputFile contains invalid values (b.yuv)" in r.stderr


def test_png_sequences(tmp_path, oracle):
    """FileFormat PNG (xSeqPNG): frames decoded to 8-bit R, G, B planes, metrics on them as they are
    (ColorSpaceInput YCbCr: no conversion) and after an RGB -> YCbCr_BT709 conversion."""
    from tests.png_writer import expected_rgb8, png_bytes, random_image
    w, h = 48, 32
    planes = {0: [], 1: []}
    for f in range(3):
        for i, colour in ((0, 2), (1, 6)):
            pix, _ = random_image(w, h, colour, 800 + 10 * f + i)
            if i == 1:  # the tested sequence: the reference frame plus a shift and noise, with alpha
                base = planes[0][f].transpose(1, 2, 0).astype(np.int64)
                pix[..., :3] = np.clip(np.roll(base, 1, axis=1) + (pix[..., :3] % 5) - 2, 0, 255).astype(np.uint8)
            (tmp_path / f"{'ab'[i]}_{f + 1:03d}.png").write_bytes(png_bytes(pix, colour, interlace=f == 1))
            planes[i].append(expected_rgb8(pix, colour))
    print("cli PNG sequences", flush=True)
    r = run_cli(["-i0", "a_{:03d}.png", "-i1", "b_{:03d}.png", "-ff", "PNG", "-ps", f"{w}x{h}",
                 "-ml", "PSNR,WSPSNR,IVPSNR", "-v", 2], tmp_path)
    prm = oracle.params(w, h, 8, 444)
    res = [oracle.frame(planes[0][f].reshape(-1), planes[1][f].reshape(-1), prm) for f in range(3)]
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR", "IVPSNR"])
    check_output(r, lines, avg)
    print("cli PNG sequences RGB -> YCbCr_BT709", flush=True)
    r = run_cli(["-i0", "a_{:03d}.png", "-i1", "b_{:03d}.png", "-ff", "PNG", "-ps", f"{w}x{h}", "-csi", "RGB",
                 "-csm", "YCbCr_BT709", "-ml", "PSNR,IVPSNR", "-v", 2], tmp_path)
    conv = lambda p: oracle.colorspace(p, 8, 1, False).reshape(-1).astype(np.uint8)
    res = [oracle.frame(conv(planes[0][f]), conv(planes[1][f]), prm) for f in range(3)]
    lines, avg = expected_psnr_family(res, ["PSNR", "IVPSNR"])
    check_output(r, lines, avg)


def _cli_values(stdout):
    """The frame and average lines of a run (the timing lines aside)."""
    return [l for l in stdout.splitlines() if l.startswith(("Frame ", "Average "))]


@pytest.mark.parametrize("i", [1, 3, 5])
def test_file_formats(tmp_path, oracle, i):
    """4:2:2 and odd-width 4:2:0 files (xSeq's unpack, libs/pmbb-core/src/xSeq.cpp:179-249)."""
    from tests.inputs import FORMAT_CASES, format_pair
    w, h, bd, fcf, seed = FORMAT_CASES[i]
    frames = [format_pair(w, h, bd, fcf, seed + k) for k in range(2)]
    (tmp_path / "a.yuv").write_bytes(b"".join(f[0].tobytes() for f in frames))
    (tmp_path / "b.yuv").write_bytes(b"".join(f[1].tobytes() for f in frames))
    print(f"cli format {w}x{h} bd{bd} cf{fcf}", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-cf", fcf, "-ml", "PSNR,WSPSNR,IVPSNR",
                 "-v", 2], tmp_path)
    prm = oracle.params(w, h, bd, fcf)
    res = [oracle.frame(a, b, prm) for a, b in frames]
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR", "IVPSNR"])
    check_output(r, lines, avg)


@pytest.mark.parametrize("action", ["WARN", "SKIP", "CNCL", "STOP"])
def test_invalid_pel_actions(tmp_path, oracle, action):
    """InvalidPelActn (xAppQMIV.cpp:681-683, :753-770): the check's report is the reference's own text
    (tests/golden/formats.json, xPicP::check); WARN and SKIP compute over the samples as read, CNCL
    over the clipped pictures, STOP ends the run."""
    import json
    import os
    from tests.inputs import RANGE_CASES, range_pair
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "formats.json")) as f:
        want = json.load(f)["range"][0]
    w, h, bd, cf, seed, inj0, inj1, _ = RANGE_CASES[0]
    a, b = range_pair(w, h, bd, cf, seed, inj0, inj1)
    write_seq(tmp_path / "a.yuv", [a], bd)
    write_seq(tmp_path / "b.yuv", [b], bd)
    print(f"cli invalid pels {action}", flush=True)
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-ml", "PSNR,WSPSNR,IVPSNR", "-ipa",
                 action, "-v", 2], tmp_path)
    report = want["check_report"].replace("{i0}", "a.yuv").replace("{i1}", "b.yuv")
    if action == "SKIP":
        assert "FILE BROKEN" not in r.stdout and "DETECTED" not in r.stdout
    else:
        assert report in r.stdout
    if action == "STOP":
        assert r.returncode != 0 and "InputFile contains invalid values (a.yuv)" in r.stderr
        return
    res = [oracle.frame(a, b, oracle.params(w, h, bd, cf, conceal=action == "CNCL"))]
    lines, avg = expected_psnr_family(res, ["PSNR", "WSPSNR", "IVPSNR"])
    check_output(r, lines, avg)


@pytest.mark.parametrize("masked", [False, True])
def test_staged_timing_report(tmp_path, masked):
    """VerboseLevel 3: the frame runs stage by stage (processor-level calls) and the application's
    AvgTime table is printed (xAppQMIV.cpp:995-1030); every frame and average line equals the fused
    run's."""
    w, h, bd = 128, 64, 10
    pairs = frames_for(w, h, bd, 420, 3, 900)
    write_seq(tmp_path / "a.yuv", [p.ref for p in pairs], bd)
    write_seq(tmp_path / "b.yuv", [p.tst for p in pairs], bd)
    args = ["-i0", "a.yuv", "-i1", "b.yuv", "-ps", f"{w}x{h}", "-bd", bd, "-erp"]
    if masked:
        write_seq(tmp_path / "m.yuv", [make_mask(w, h, 8, 900 + i, kind="soft") for i in range(3)], 8)
        args += ["-im", "m.yuv", "-bdm", 8, "-ml", "PSNR,WSPSNR,IVPSNR"]
        names = ["PSNR", "WSPSNR", "IVPSNR"]
    else:
        args += ["-ml", "All"]
        names = ["PSNR", "WSPSNR", "IVPSNR", "SSIM", "MSSSIM", "IVSSIM"]
    print(f"cli staged timing masked={masked}", flush=True)
    fused = run_cli(args + ["-v", 2], tmp_path)
    staged = run_cli(args + ["-v", 3], tmp_path)
    assert fused.returncode == 0 and staged.returncode == 0, staged.stdout[-2000:] + staged.stderr
    assert _cli_values(staged.stdout) == _cli_values(fused.stdout)
    out = staged.stdout.splitlines()
    for stage in ("LOAD", "VALIDATE", "PREPROC", "GCD") + (() if masked else ("SCP",)):
        assert any(l.startswith("AvgTime ") and l.split()[1] == stage and l.endswith(" ms") for l in out), stage
    for name in names:
        line = [l for l in out if l.startswith("AvgTime ") and l.split()[1] == name]
        assert len(line) == 1 and " ms" in line[0], name
        if name in ("IVPSNR", "IVSSIM"):
            assert "   Total " in line[0]
