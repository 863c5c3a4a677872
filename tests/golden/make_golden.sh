#!/bin/bash
# Regenerates the golden fixtures in this folder from the SPADE reference.
# Runs only where the reference sources are present (/root/reference): builds
# oracle/ref/ref_driver.cpp against the reference headers in place (output in
# oracle/_ref/) and writes the raw float64 fixtures plus manifest.txt here.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
root="$(cd "$here/../.." && pwd)"
make -C "$root/oracle/ref" REF=/root/reference/src
"$root/oracle/_ref/ref_driver" "$here"
( cd "$here" && sha256sum *.bin vtk_g1/*.pvts vtk_g1/data_g1/*.vts > SHA256SUMS )
# the compiled reference's own single-core flux_div time on the 64^3 TGV
# (4^3 blocks of 16^3, cent_keep<2> + visc_lr power law), on this host
cpu="$(grep -m1 'model name' /proc/cpuinfo | cut -d: -f2 | sed 's/^ *//')"
rate="$("$root/oracle/_ref/ref_driver" --bench 4 5 | tail -1 | sed 's/.*"mcell_rhs_per_s": \([0-9.]*\).*/\1/')"
echo "timing c0_flux_div_one_core_mcell_rhs_per_s $rate ($cpu)" >> "$here/manifest.txt"
