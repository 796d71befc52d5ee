"""Makes tests/golden/settings/: the reference's run-configuration files (settings/*.ini of
mdelorme/fv2d), copied unchanged as input fixtures of the .ini reader and driver tests.

    python tests/golden/copy_settings.py [REFERENCE_ROOT]   (default /root/reference)

They are data (sections and key = value lines), not source. The tests read them in place;
this script is only how they were made and can re-check that they still match.
"""
import filecmp
import os
import shutil
import sys

NAMES = ["C91.ini", "H84_chi1.5.ini", "blast.ini", "diffusion.ini", "gresho_vortex.ini", "kelvin_helmholtz.ini",
         "rayleigh_taylor.ini", "sod_x.ini", "sod_y.ini"]


def main(argv):
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "settings")
    os.makedirs(out, exist_ok=True)
    for name in NAMES:
        src = os.path.join(ref, "settings", name)
        dst = os.path.join(out, name)
        if os.path.exists(dst) and filecmp.cmp(src, dst, shallow=False):
            print(f"{name}: unchanged")
            continue
        shutil.copyfile(src, dst)
        print(f"{name}: copied")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
