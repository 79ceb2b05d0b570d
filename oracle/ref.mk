# Builds the reference's serial Fortran model (original/nemolite2d.f90) and
# the two support libraries it links (dl_esm_inf field/log helpers and
# dl_timer) directly from their source files under /root/reference, with
# AMD flang and gcc.  Every output (objects, .mod files, the executable)
# goes to oracle/_ref/.  The reference's own Makefiles are not used.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags: IEEE double arithmetic (-O2, no fast-math, no FMA contraction), so
# that the binary's results are reproducible; the reference's own gnu flags
# are -Ofast -march=core2 (compiler_setup/gnu.sh:10-11).  A second build of
# the model, nemolite2d_ref_fast (objects in $(OUT)/fast), uses flang's
# equivalent, -O3 -ffast-math -fstack-arrays (what flang's deprecated -Ofast
# means; its startup sets flush-to-zero), for an AVX2/FMA x86-64 target: the
# CPU baseline as the reference's users would build it.
REF     ?= /root/reference
OUT     ?= oracle/_ref
FLANG   ?= flang
GCC     ?= gcc
FFLAGS  ?= -O2 -ffp-contract=off -fopenmp
CFLAGS  ?= -O2

INF  = $(REF)/shared/dl_esm_inf/finite_difference/src
TIM  = $(REF)/shared/dl_timer/src
ORIG = $(REF)/original

INF_SRC = kind_params_mod.f90 global_parameters_mod.F90 region_mod.f90 \
          gocean_mod.F90 halo_mod.f90 grid_mod.f90 field_mod.f90
TIM_F   = dl_timer_constants_mod.f90 intel_timer_mod.f90 parallel_stub.f90 dl_timer.f90
TIM_C   = posix_clock.c gettofd.c intel_timer.c

INF_OBJ = $(addprefix $(OUT)/inf_,$(addsuffix .o,$(basename $(INF_SRC))))
TIM_FOBJ = $(addprefix $(OUT)/tim_,$(addsuffix .o,$(basename $(TIM_F))))
TIM_COBJ = $(addprefix $(OUT)/timc_,$(addsuffix .o,$(basename $(TIM_C))))

EXE = $(OUT)/nemolite2d_ref
FAST_DIR   = $(OUT)/fast
FAST_EXE   = $(OUT)/nemolite2d_ref_fast
FFLAGS_FAST ?= -O3 -ffast-math -fstack-arrays -march=x86-64-v3 -fopenmp

# The same model with one added statement: after each step of the time loop
# (original/nemolite2d.f90:74-77) it prints the step's checksums
# SUM(ABS(.)) of sshn, un and vn over (1:jpi,1:jpj) -- the reference's own
# field_checksum (shared/dl_esm_inf/finite_difference/src/field_mod.f90:
# 935-956) -- to 17 significant digits (ES24.16).  The scratch copy the
# statement is added to lives in $(OUT) only.
TRACE_SRC = $(OUT)/nemolite2d_trace.f90
TRACE_EXE = $(OUT)/nemolite2d_ref_trace

all: $(EXE) $(TRACE_EXE) $(FAST_EXE)

$(OUT):
	mkdir -p $(OUT)

# The module dependency order is the order of the lists above, so objects
# are built serially in that order.
$(EXE): | $(OUT)
	set -e; for f in $(INF_SRC); do \
	  $(FLANG) $(FFLAGS) -module-dir $(OUT) -I$(OUT) -c $(INF)/$$f -o $(OUT)/inf_$${f%.*}.o; done
	set -e; for f in $(TIM_F); do \
	  $(FLANG) $(FFLAGS) -module-dir $(OUT) -I$(OUT) -c $(TIM)/$$f -o $(OUT)/tim_$${f%.*}.o; done
	set -e; for f in $(TIM_C); do \
	  $(GCC) $(CFLAGS) -c $(TIM)/$$f -o $(OUT)/timc_$${f%.*}.o; done
	$(FLANG) $(FFLAGS) -module-dir $(OUT) -I$(OUT) -c $(ORIG)/nemolite2d.f90 -o $(OUT)/nemolite2d.o
	$(FLANG) $(FFLAGS) -o $@ $(OUT)/nemolite2d.o $(INF_OBJ) $(TIM_FOBJ) $(TIM_COBJ)

$(TRACE_EXE): $(EXE) $(ORIG)/nemolite2d.f90
	sed -e '/^ *CALL step *$$/a\
        WRITE (*, "(A,I8,3ES24.16)") "step checksums ", istp, SUM(ABS(sshn(1:jpi, 1:jpj))), \&\
            SUM(ABS(un(1:jpi, 1:jpj))), SUM(ABS(vn(1:jpi, 1:jpj)))' \
	    $(ORIG)/nemolite2d.f90 > $(TRACE_SRC)
	test "$$(grep -c '"step checksums "' $(TRACE_SRC))" = 1
	mkdir -p $(OUT)/trace_mod
	$(FLANG) $(FFLAGS) -module-dir $(OUT)/trace_mod -I$(OUT) -c $(TRACE_SRC) -o $(OUT)/nemolite2d_trace.o
	$(FLANG) $(FFLAGS) -o $@ $(OUT)/nemolite2d_trace.o $(INF_OBJ) $(TIM_FOBJ) $(TIM_COBJ)

$(FAST_EXE): $(EXE)
	mkdir -p $(FAST_DIR)
	set -e; for f in $(INF_SRC); do \
	  $(FLANG) $(FFLAGS_FAST) -module-dir $(FAST_DIR) -I$(FAST_DIR) -c $(INF)/$$f -o $(FAST_DIR)/inf_$${f%.*}.o; done
	set -e; for f in $(TIM_F); do \
	  $(FLANG) $(FFLAGS_FAST) -module-dir $(FAST_DIR) -I$(FAST_DIR) -c $(TIM)/$$f -o $(FAST_DIR)/tim_$${f%.*}.o; done
	$(FLANG) $(FFLAGS_FAST) -module-dir $(FAST_DIR) -I$(FAST_DIR) -c $(ORIG)/nemolite2d.f90 -o $(FAST_DIR)/nemolite2d.o
	$(FLANG) $(FFLAGS_FAST) -o $@ $(FAST_DIR)/nemolite2d.o $(addprefix $(FAST_DIR)/,$(notdir $(INF_OBJ) $(TIM_FOBJ))) \
	    $(TIM_COBJ)

clean:
	rm -rf $(OUT)

.PHONY: all clean
