"""Copies the part of the reference's sample data that the real-data tests
use into tests/golden/sample_data (data files only: the wrist calibrations,
the four wrist bone volumes, one radiograph frame per wrist camera and the
two filter settings files), and writes wrist.cfg, the GUI's built-in wrist
trial (AutoscoperMainWindow.cpp:2189-2244) as a version 1.1 file with
paths relative to it.

    python tests/golden/copy_sample_data.py [/path/to/reference]
"""
import shutil
import sys
from pathlib import Path

FILES = [
    "Calibration/xr_calib_wrist_cam01.txt",
    "Calibration/xr_calib_wrist_cam02.txt",
    "Models/rad_dcm_cropped.tif",
    "Models/mc2_mc3_dcm_cropped.tif",
    "Models/mc3_dcm_cropped.tif",
    "Models/uln_dcm_cropped.tif",
    "XMA_UND/xr_data_wrist_cam01/xr_rad_uln_cam01_UND.0099.tif",
    "XMA_UND/xr_data_wrist_cam02/xr_rad_uln_cam02_UND.0099.tif",
    "xParameters/control_settings.vie",
    "xParameters/empty_filter.vie",
]

CFG = """# The wrist sample trial (radius, MC2-MC3, MC3, ulna), one frame per camera
Version 1.1

mayaCam_csv Calibration/xr_calib_wrist_cam01.txt
mayaCam_csv Calibration/xr_calib_wrist_cam02.txt

CameraRootDir XMA_UND/xr_data_wrist_cam01
CameraRootDir XMA_UND/xr_data_wrist_cam02

VolumeFile Models/rad_dcm_cropped.tif
VolumeFlip 0 0 0
VoxelSize 0.39625 0.39625 0.625
VolumeFile Models/mc2_mc3_dcm_cropped.tif
VolumeFlip 0 0 0
VoxelSize 0.39625 0.39625 0.625
VolumeFile Models/mc3_dcm_cropped.tif
VolumeFlip 0 0 0
VoxelSize 0.39625 0.39625 0.625
VolumeFile Models/uln_dcm_cropped.tif
VolumeFlip 0 0 0
VoxelSize 0.39625 0.39625 0.625

RenderResolution 512 512
OptimizationOffsets 0.1 0.1 0.1 0.1 0.1 0.1
"""


def main(ref: Path) -> None:
    out = Path(__file__).resolve().parent / "sample_data"
    for rel in FILES:
        dst = out / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(ref / "sample_data" / rel, dst)
    (out / "wrist.cfg").write_text(CFG)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
