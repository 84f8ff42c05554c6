# Builds the reference Fortran solver out of tree, from the sources where
# they lie, into oracle/_ref/ (git-ignored).  Used only to generate and
# check the golden fixtures and as the CPU baseline; never shipped.
#
#   make -f oracle/ref.mk            # fixture build + timing build + driver
#
# Toolchain: LLVM flang from ROCm and the MPICH 3.3.2 installed in
# /opt/conda (headers + libmpifort/libmpi).  The fixture build disables
# FMA contraction so the arithmetic is plain IEEE double, which the
# numpy/C restatements reproduce term by term.
REF      ?= /root/reference
HERE     := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT      := $(HERE)_ref
FC       := $(shell command -v flang || echo /opt/rocm/bin/amdflang)
MPIINC   := /opt/conda/include
MPILIB   := -L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmpifort -lmpi
FIXFLAGS := -O2 -ffp-contract=off
TIMFLAGS := -O3

# makefile order of the reference (makefile:7-13); setup2d first
UNITS := setup2d readinput readGRIDinfo connectivity connect_bdry \
	map_proc_match map_cyc_match_rem getIVCELL_proc getIVCELL_cyc \
	map_cyc_loc initsetup mapder calcjacob setinitialcond \
	tecplotter2dsetup tecplotter compflux getrusanov interfaceflux_all \
	procintflux_all cycremflux_all cyclocflux_all compresid iterations \
	write_data read_data_restart bcflux AV

FIXOBJ := $(UNITS:%=$(OUT)/fix/%.o)
TIMOBJ := $(UNITS:%=$(OUT)/tim/%.o)

all: $(OUT)/sd2d_MHD $(OUT)/sd2d_MHD_O3 $(OUT)/ref_driver

# The reference program with its compresid replaced by the MI355X library
# (sddc_amd/fortran): reference objects minus compresid.o, plus the
# ISO_C_BINDING module and the drop-in compresid, linked to libsd2d.so.
# Needs sddc_amd/lib/libsd2d.so built first.
LIBDIR   := $(abspath $(HERE)../sddc_amd/lib)
FDIR     := $(abspath $(HERE)../sddc_amd/fortran)
DROPOBJ  := $(filter-out $(OUT)/fix/compresid.o,$(FIXOBJ))
dropin: $(OUT)/sd2d_MHD_dropin

$(OUT)/sd2d_MHD_dropin: $(OUT)/fix/.stamp $(FDIR)/sd2d_c.f90 $(FDIR)/compresid_sd2d.f90 $(LIBDIR)/libsd2d.so
	mkdir -p $(OUT)/drop
	$(FC) -c $(FIXFLAGS) -module-dir $(OUT)/drop $(FDIR)/sd2d_c.f90 -o $(OUT)/drop/sd2d_c.o
	$(FC) -c $(FIXFLAGS) -I$(MPIINC) -I$(OUT)/fix -I$(OUT)/drop -module-dir $(OUT)/drop \
	  $(FDIR)/compresid_sd2d.f90 -o $(OUT)/drop/compresid_sd2d.o
	$(FC) $(FIXFLAGS) -o $@ $(OUT)/fix/main.o $(DROPOBJ) $(OUT)/drop/sd2d_c.o \
	  $(OUT)/drop/compresid_sd2d.o -L$(LIBDIR) -lsd2d \
	  -Wl,-rpath,'$$ORIGIN/../../sddc_amd/lib' -Wl,-rpath,/opt/rocm/lib $(MPILIB)

$(OUT)/fix $(OUT)/tim:
	mkdir -p $@

# units compile serially in makefile order (module dependencies)
$(OUT)/fix/.stamp: | $(OUT)/fix
	set -e; for u in $(UNITS); do \
	  $(FC) -c $(FIXFLAGS) -I$(MPIINC) -module-dir $(OUT)/fix \
	    $(REF)/$$u.f90 -o $(OUT)/fix/$$u.o; done; \
	$(FC) -c $(FIXFLAGS) -I$(MPIINC) -I$(OUT)/fix -module-dir $(OUT)/fix \
	  $(REF)/main.f90 -o $(OUT)/fix/main.o; touch $@

$(OUT)/tim/.stamp: | $(OUT)/tim
	set -e; for u in $(UNITS); do \
	  $(FC) -c $(TIMFLAGS) -I$(MPIINC) -module-dir $(OUT)/tim \
	    $(REF)/$$u.f90 -o $(OUT)/tim/$$u.o; done; \
	$(FC) -c $(TIMFLAGS) -I$(MPIINC) -I$(OUT)/tim -module-dir $(OUT)/tim \
	  $(REF)/main.f90 -o $(OUT)/tim/main.o; touch $@

$(OUT)/sd2d_MHD: $(OUT)/fix/.stamp
	$(FC) $(FIXFLAGS) -o $@ $(OUT)/fix/main.o $(FIXOBJ) $(MPILIB)

$(OUT)/sd2d_MHD_O3: $(OUT)/tim/.stamp
	$(FC) $(TIMFLAGS) -o $@ $(OUT)/tim/main.o $(TIMOBJ) $(MPILIB)

$(OUT)/ref_driver: $(OUT)/fix/.stamp $(HERE)ref_driver.f90
	$(FC) -c $(FIXFLAGS) -I$(MPIINC) -I$(OUT)/fix -module-dir $(OUT)/fix \
	  $(HERE)ref_driver.f90 -o $(OUT)/fix/ref_driver.o
	$(FC) $(FIXFLAGS) -o $@ $(OUT)/fix/ref_driver.o $(FIXOBJ) $(MPILIB)

clean:
	rm -rf $(OUT)

.PHONY: all clean dropin
