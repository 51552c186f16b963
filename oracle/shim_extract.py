"""Writes the reference-side binding shown in INTEGRATION.md (section 2, the C++ block a
maintainer adds next to PIMCPass in mc_main.cc) into a compilable file, with the includes
mc_main.cc has (mc_main.cc:5-26) and one exported table of the shim's functions, so that a
link with --no-undefined must resolve everything they use.  Test infrastructure: run by the
`shim` target of oracle/ref.mk in the build container, where the reference sources are."""
import re
import sys

HEAD = """// extracted from INTEGRATION.md section 2 by oracle/shim_extract.py
#include <stdlib.h>
#include <sstream>
#include <iomanip>
#include <cmath>
#include "mc_confg.h"
#include "mc_setup.h"
#include "mc_input.h"
#include "mc_randg.h"
#include "mc_utils.h"
#include "mc_poten.h"
#include "mc_piqmc.h"
#include "mc_qworm.h"
#include "mc_estim.h"
#include "mc_const.h"
#include <stdio.h>
#include <omp.h>
#include "rngstream.h"
#include "omprng.h"
"""

TAIL = """
// every function of the shim, so that the linker has to resolve what they call
extern "C" {
void (*moribs_shim_entries[3])() = {gpu_init, sync_to_gpu, sync_from_gpu};
}
"""


def main(src, dst):
    text = open(src).read()
    sec = text[text.index("## 2."):]
    sec = sec[:sec.index("\n## 3.")]
    blocks = re.findall(r"```cpp\n(.*?)```", sec, flags=re.S)
    if not blocks:
        raise SystemExit("no C++ block in section 2 of " + src)
    with open(dst, "w") as f:
        f.write(HEAD + blocks[0] + TAIL)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
