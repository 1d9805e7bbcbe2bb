#!/bin/bash
# Builds the reference pyqg kernel (kernel.pyx, numpy.fft fallback: pyfftw is not
# installable offline) from its source where it lies under /root/reference, into
# tests/golden/_refbuild/ (git-ignored and never uploaded to a GPU host).
# Used only by make_golden.py, in the development container, to write the golden
# vectors committed beside this script.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="$here/_refbuild"
mkdir -p "$out"
# kernel.pyx includes '.compile_time_use_pyfftw.pxi' (written by the reference's
# setup.py); supply it through the include path instead of touching the read-only tree.
printf 'DEF PYQG_USE_PYFFTW = 0\n' > "$out/.compile_time_use_pyfftw.pxi"
python3 -m cython -3 -I "$out" -o "$out/kernel.c" /root/reference/pyqg/kernel.pyx
ext=$(python3 -c "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))")
gcc -shared -fPIC -O2 -fopenmp \
    $(python3-config --includes) -I"$(python3 -c 'import numpy; print(numpy.get_include())')" \
    "$out/kernel.c" -o "$out/kernel$ext"
echo "built $out/kernel$ext"
