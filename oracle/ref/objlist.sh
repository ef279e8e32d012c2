#!/bin/sh
# Prints the reference's serial object list (module/file.cpp), read from the
# object-name lists in its makefiles, without the CUDA and OpenCL backends.
REF=${1:-/root/reference}
cd "$REF" || exit 1
{
  for f in src/*/makefile; do
    d=$(basename "$(dirname "$f")")
    case $d in cuda|opencl) continue ;; esac
    awk -v d="$d" '/_objects *=/{on=1} on{for(i=1;i<=NF;i++){t=$i; sub(/\\$/,"",t); if(t ~ /\.o$/) print d"/"t}} on && !/\\$/{on=0}' "$f"
  done
  grep -o 'obj/[a-z]*/[A-Za-z0-9_]*\.o' makefile | sed 's/^obj\///' | grep -v -e '^cuda' -e '^opencl'
} | sort -u | sed 's/\.o$/.cpp/'
