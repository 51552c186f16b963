# Builds the upstream MoRiBS-PIMC sources, in place under REF (read-only), into
# oracle/_ref/.  Test infrastructure only: nothing in the product links this.
#
#   make -f oracle/ref.mk            -> oracle/_ref/pimc          (the reference program)
#                                       oracle/_ref/libref.so     (reference objects + harness)
#
# Two flavours of the reference objects are built:
#   * _ref/obj_fast : the reference's own optimisation flags (-Ofast, OpenMP), with
#                     -march=x86-64-v3 instead of -march=native so the binary runs on
#                     any AVX2 host; used for the CPU-baseline timing.
#   * _ref/obj_gold : -O2, no fast-math; used for golden vectors (bit-stable).
# SPRNG's 48-bit LCG (the generator the reference links, -llcg) is compiled from the
# vendored C sources; no prebuilt library or binary from the reference is used.

REF     ?= /root/reference
OUT     := $(dir $(lastword $(MAKEFILE_LIST)))_ref
FC      := /opt/rocm/lib/llvm/bin/flang
FRT     := $(firstword $(wildcard /opt/rocm/lib/llvm/lib/clang/*/lib/x86_64-unknown-linux-gnu/libflang_rt.runtime.a))
CXX     := g++
CC      := gcc

FAST    := -Ofast -march=x86-64-v3 -fopenmp
GOLD    := -O2 -fopenmp -fno-fast-math
DEFS    := -DMOVECENTROIDTEST -I$(REF) -I$(REF)/sprng/include -I$(REF)/sprng/SRC

CXXSRC  := mc_piqmc mc_estim mc_qworm mc_input mc_setup mc_poten mc_randg mc_utils omprng rngstream
F77SRC  := rotden rotpro_sub rotred potred vcord vcalc initconf vspher caleng_tip4p_gg
SPRNGC  := fwrap_mpi cputime makeseed simple_mpi memory communicate checkid store
SPRNGL  := lcg primes

FASTO   := $(addprefix $(OUT)/obj_fast/,$(addsuffix .o,$(CXXSRC) $(F77SRC) $(SPRNGC) $(SPRNGL)))
GOLDO   := $(addprefix $(OUT)/obj_gold/,$(addsuffix .o,$(CXXSRC) $(F77SRC) $(SPRNGC) $(SPRNGL)))

.PHONY: all
all: $(OUT)/pimc $(OUT)/libref.so

$(OUT)/obj_fast $(OUT)/obj_gold:
	mkdir -p $@

# --- C++ sources of the reference -------------------------------------------
$(OUT)/obj_fast/%.o: $(REF)/%.cc | $(OUT)/obj_fast
	$(CXX) $(FAST) -fPIC $(DEFS) -w -c $< -o $@
$(OUT)/obj_gold/%.o: $(REF)/%.cc | $(OUT)/obj_gold
	$(CXX) $(GOLD) -fPIC $(DEFS) -w -c $< -o $@

# --- Fortran 77 sources of the reference ------------------------------------
$(OUT)/obj_fast/%.o: $(REF)/%.f | $(OUT)/obj_fast
	$(FC) -O3 -fPIC -c $< -o $@
$(OUT)/obj_gold/%.o: $(REF)/%.f | $(OUT)/obj_gold
	$(FC) -O2 -fPIC -c $< -o $@

# --- SPRNG (lcg) from the vendored C sources --------------------------------
SPRNGFLAGS := -O2 -fPIC -DLittleEndian -DLINUX -DAdd_ -I$(REF)/sprng/SRC -w
$(OUT)/obj_fast/%.o: $(REF)/sprng/SRC/%.c | $(OUT)/obj_fast
	$(CC) $(SPRNGFLAGS) -c $< -o $@
$(OUT)/obj_gold/%.o: $(REF)/sprng/SRC/%.c | $(OUT)/obj_gold
	$(CC) $(SPRNGFLAGS) -c $< -o $@
$(OUT)/obj_fast/%.o: $(REF)/sprng/SRC/lcg/%.c | $(OUT)/obj_fast
	$(CC) $(SPRNGFLAGS) -c $< -o $@
$(OUT)/obj_gold/%.o: $(REF)/sprng/SRC/lcg/%.c | $(OUT)/obj_gold
	$(CC) $(SPRNGFLAGS) -c $< -o $@

# --- the reference program (CPU baseline) -----------------------------------
$(OUT)/obj_fast/mc_main.o: $(REF)/mc_main.cc | $(OUT)/obj_fast
	$(CXX) $(FAST) -fPIC $(DEFS) -w -c $< -o $@

$(OUT)/pimc: $(OUT)/obj_fast/mc_main.o $(FASTO)
	$(CXX) -fopenmp -o $@ $^ $(FRT) -lm

# --- golden-vector harness: reference objects + oracle/ref_harness.cc -------
# mc_main.cc is compiled with its main() renamed so PIMCPass/MCGetAverage link.
$(OUT)/obj_gold/mc_main.o: $(REF)/mc_main.cc | $(OUT)/obj_gold
	$(CXX) $(GOLD) -fPIC $(DEFS) -Dmain=moribs_reference_main -w -c $< -o $@

$(OUT)/obj_gold/ref_harness.o: $(dir $(lastword $(MAKEFILE_LIST)))ref_harness.cc | $(OUT)/obj_gold
	$(CXX) $(GOLD) -fPIC $(DEFS) -w -c $< -o $@

# the harness supplies the RNG entry points, so SPRNG and omprng.o are left out
HARNO := $(addprefix $(OUT)/obj_gold/,$(addsuffix .o,$(filter-out omprng,$(CXXSRC)) $(F77SRC)))
$(OUT)/libref.so: $(OUT)/obj_gold/ref_harness.o $(OUT)/obj_gold/mc_main.o $(HARNO)
	$(CXX) -shared -fopenmp -o $@ $^ $(FRT) -lm

# --- link check of the reference-side binding shown in INTEGRATION.md (section 2) ----------
# The shim is extracted from the document, compiled against the reference headers and linked
# with the reference objects and the product library into a shared object with no undefined
# symbols: the documented binding builds as written.
ODIR    := $(dir $(lastword $(MAKEFILE_LIST)))
PKGDIR  := $(abspath $(ODIR)../moribs-pimc_amd)
$(OUT)/shim:
	mkdir -p $@
$(OUT)/shim/gpu_shim.cc: $(ODIR)../INTEGRATION.md $(ODIR)shim_extract.py | $(OUT)/shim
	python3 $(ODIR)shim_extract.py $< $@
$(OUT)/shim/gpu_shim.o: $(OUT)/shim/gpu_shim.cc $(ODIR)../include/moribs_pimc.h
	$(CXX) $(GOLD) -fPIC $(DEFS) -I$(ODIR)../include -Wall -Wno-unused-function -c $< -o $@
$(OUT)/libgpu_shim.so: $(OUT)/shim/gpu_shim.o $(OUT)/obj_gold/mc_main.o $(GOLDO)
	$(CXX) -shared -fopenmp -o $@ $^ $(FRT) -lm -L$(PKGDIR) -lmoribs_pimc -Wl,-rpath,$(PKGDIR) -Wl,--no-undefined

.PHONY: shim
shim: $(OUT)/libgpu_shim.so
