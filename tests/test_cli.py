"""The command-line driver without a GPU: parameter parsing (command line and configuration files),
the configuration report, frame counting and the reference application's error messages
(apps/QMIV/src/xAppQMIV.cpp:205-339, :341-392, :492-560). Runs stop at device initialisation,
which needs a GPU."""
import os

import numpy as np
import pytest

from tests.cli_format import CLI, run_cli, write_seq


@pytest.fixture(scope="module")
def seqs(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    w, h = 16, 8
    frame = np.zeros(w * h * 3 // 2, np.uint16)
    write_seq(d / "a.yuv", [frame] * 5, 10)
    write_seq(d / "b.yuv", [frame] * 3, 10)
    with open(d / "b.yuv", "ab") as f:
        f.write(b"\0" * 17)  # a partial frame at the end is not counted
    write_seq(d / "m.yuv", [frame] * 2, 10)
    return d


def config_block(out):
    lines = out.splitlines()
    beg = lines.index("Run-time configuration:")
    return {k.strip(): v.strip() for k, v in (l.split("=", 1) for l in lines[beg + 1:] if "=" in l)}


def test_binary_is_built():
    assert os.access(CLI, os.X_OK)


def test_configuration_report_and_frame_count(seqs):
    r = run_cli(["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-bd", "10", "-s0", "3", "-erp", "-sr", "3",
                 "-cwa", "2:1:1:0", "-ml", "psnr, IVPSNR,SSIM", "-v", "1"], seqs)
    assert r.returncode != 0  # no device here
    cfg = config_block(r.stdout)
    assert cfg["InputFile0"] == "a.yuv" and cfg["PictureSize"] == "16x8" and cfg["BitDepth"] == "10"
    assert cfg["ChromaFormat"] == "CF420" and cfg["StartFrame0"] == "3" and cfg["NumberOfFrames"] == "-1  (all)"
    assert cfg["MetricList"] == "PSNR, IVPSNR, SSIM"
    assert cfg["Equirectangular"] == "1" and cfg["LonRangeDeg"] == "360"
    assert cfg["SearchRange"] == "3  (custom)" and cfg["CmpWeightsSearch"] == "4:1:1:0  (default)"
    assert cfg["CmpWeightsAverage"] == "2:1:1:0  (custom)" and cfg["UnnoticeableCoef"] == "0.01:0.01:0.01:0.00  (default)"
    assert cfg["WindowSize"] == "7x7" and cfg["PictureMargin"] == "4" and cfg["UseMask"] == "0"
    assert cfg["ColorSpaceInput"] == "YCbCr  (irrelevant)"
    assert cfg["DetectedFrames0"] == "5" and cfg["DetectedFrames1"] == "3"
    # min(min(5, 3), min(5 - 3, 3 - 0)) (xAppQMIV.cpp:553-555)
    assert cfg["FramesToProcess"] == "2"
    assert "ERROR --> device initialisation" in r.stderr


def test_configuration_file_and_override(seqs):
    (seqs / "run.cfg").write_text("# comment\nInputFile0 = a.yuv\nInputFile1 = \"b.yuv\"\nPictureWidth = 16\n"
                                  "PictureHeight = 8\nPictureFormat = yuv420p10le\nNumberOfFrames = 1\n"
                                  "InputFileM = m.yuv\nBitDepthM = 10\nColorSpaceInput = YCbCr_BT709\n"
                                  "ColorSpaceMetric = RGB\nMetricList = PSNR, WSPSNR\n")
    r = run_cli(["-c", "run.cfg", "-nf", "2", "-v", "1"], seqs)
    cfg = config_block(r.stdout)
    assert cfg["PictureSize"] == "16x8" and cfg["BitDepth"] == "10" and cfg["NumberOfFrames"] == "2"
    assert cfg["InputFileM"] == "m.yuv" and cfg["BitDepthM"] == "10" and cfg["UseMask"] == "1"
    assert cfg["ColorSpaceInput"] == "YCbCr_BT709" and cfg["ColorSpaceMetric"] == "RGB"
    assert cfg["DetectedFrames2"] == "2" and cfg["FramesToProcess"] == "2"
    assert "SizeOfInputFileM = 768" in r.stdout


def test_png_sequences_are_counted(seqs):
    from tests.png_writer import png_bytes
    for i in range(1, 4):
        (seqs / f"a{i}.png").write_bytes(png_bytes(np.zeros((8, 16, 3), np.uint8), 2))
    for i in range(0, 6):
        (seqs / f"b{i:02d}.png").write_bytes(png_bytes(np.zeros((8, 16, 3), np.uint8), 2))
    r = run_cli(["-i0", "a{}.png", "-i1", "b{:02d}.png", "-ff", "PNG", "-ps", "16x8", "-s1", "1"], seqs)
    cfg = config_block(r.stdout)
    assert cfg["FileFormat"] == "PNG" and "SizeOfInputFile0" not in cfg
    # a: files 1..3; b: files 0..5 start at 1 (the last of 0 and 1 that exists), so 5 frames
    assert cfg["DetectedFrames0"] == "3" and cfg["DetectedFrames1"] == "5" and cfg["FramesToProcess"] == "3"


@pytest.mark.parametrize("args,message", [
    (["-i0", "a.yuv", "-ps", "16x8"], "InputFile1 is empty"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-ml", "PSNR,XYZ"], 'Token "XYZ" in not a metric'),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-bd", "16"], "Invalid or unsuported BitDepth value"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-cf", "411"], "Invalid or unsuported ChromaFormat value"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-sr", "0"], "SearchRange value must non-zero"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-csi", "YCbCr_BT709", "-csm", "YCbCr_BT2020"],
     "YCbCr to YCbCr conversion is not supported."),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-csi", "YCbCr", "-csm", "RGB"],
     "Generic YCbCr cannot be used for colorspace conversion."),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-im", "m.yuv", "-ml", "PSNR,IVSSIM"],
     "Structural Similarity metrics cannot be combined with Mask mode"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-ipa", "HALT"], "Invalid action HALT"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-ff", "TIFF"], "Invalid FileFormat TIFF"),
    (["-i0", "a{}.png", "-i1", "b{}.png", "-ps", "16x8", "-ff", "PNG", "-bd", "10"], "PNG input is 8-bit"),
    (["-i0", "none{}.png", "-i1", "b{}.png", "-ps", "16x8", "-ff", "PNG"], "InputFile0 does not exist (none{}.png)"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-xyz", "1"], "Unknown parameter -xyz"),
    (["-i0", "a.yuv", "-i1", "nope.yuv", "-ps", "16x8"], "InputFile1 does not exist (nope.yuv)"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-bd", "10", "-s1", "3"],
     "StartFrame1 >= DetectedFrames1 for (b.yuv)"),
    (["-i0", "a.yuv", "-i1", "b.yuv", "-ps", "16x8", "-bd", "10", "-im", "m.yuv", "-bdm", "10", "-ml", "PSNR"],
     "FramesToProcess > NumOfFramesM"),
])
def test_errors(seqs, args, message):
    r = run_cli(args, seqs)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr
