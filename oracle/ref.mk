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
# are -Ofast -march=core2 (compiler_setup/gnu.sh:10-11).
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

all: $(EXE)

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

clean:
	rm -rf $(OUT)

.PHONY: all clean
