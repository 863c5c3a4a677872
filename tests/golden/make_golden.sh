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
