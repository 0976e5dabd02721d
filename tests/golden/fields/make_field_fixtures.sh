#!/bin/bash
# Regenerates the field-file fixtures in this folder with a Fortran runtime (amdflang).
set -e
cd "$(dirname "$0")"
tmp=$(mktemp -d)
/opt/rocm/bin/amdflang -O0 field_writer.f90 -module-dir "$tmp" -o "$tmp/field_writer"
"$tmp/field_writer"
rm -rf "$tmp"
ls -l *.res
