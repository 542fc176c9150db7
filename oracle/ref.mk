# Builds the reference SNAPI C++ sources (read in place under REF, never copied)
# into a shared library plus a small driver, with every output under oracle/_ref/.
# The reference's own autotools build is not used: each translation unit is
# compiled directly with g++.  cmlib/fileio.cc and cmlib/fits_read.cc need
# cfitsio, which is not installed; they are left out, so the library keeps one
# undefined symbol (read_file, used only by the 3-D SPINOR cube reader, which
# the driver never calls).
#
#   make -f oracle/ref.mk            strict build (-O2, no fast-math): golden vectors
#   make -f oracle/ref.mk FAST=1     reference flags (-O3 -ffast-math): drift study
REF      ?= /root/reference
HERE     := $(dir $(lastword $(MAKEFILE_LIST)))
OUT      := $(HERE)_ref$(if $(FAST),_fast,)
OBJ      := $(OUT)/obj
CXX      ?= g++
OPT      := $(if $(FAST),-O3 -ffast-math,-O2)
CXXFLAGS := $(OPT) -fPIC -fsigned-char -w -std=gnu++98 -I$(REF) -I$(REF)/cmlib -I$(REF)/grid \
            -I$(REF)/atmos -I$(REF)/atmos/atmol -I$(REF)/atmos/atmol/atom \
            -I$(REF)/atmos/atmol/mol -I$(REF)/atmos/plpar -I$(REF)/jsub

SRCS := $(addprefix atmos/,acfg.cc atmos.cc atmos_chm.cc atmos_fio.cc atmos_fit.cc atmos_model.cc \
          atmos_obs.cc atmos_pop.cc atmos_rad.cc atmos_rad_synthesis.cc atmos_responses.cc atmos_rts.cc) \
        $(addprefix atmos/plpar/,atmos_pp.cc atmos_pp_rts.cc atmos_ppbez.cc atmos_ppbez_rts.cc) \
        atmos/atmol/atmol.cc \
        $(addprefix atmos/atmol/atom/,atom.cc atom_pop.cc atom_rad.cc atom_rad_synthesis.cc \
          atom_responses.cc atomcfg.cc bfcs.cc colr.cc partf.cc) \
        atmos/atmol/atom/H/hydrogen.cc atmos/atmol/atom/He/helium.cc \
        $(addprefix atmos/atmol/mol/,mol.cc molcfg.cc moleqc.cc) \
        grid/grid.cc grid/node.cc \
        $(addprefix cmlib/,mathtools.cc mem.cc profiles.cc obs.cc io.cc pack.cc uts.cc struts.cc \
          cfguts.cc flags.cc compress.cc net.cc cmdln.cc jnfo.cc fdset.cc ana_read.cc ana_write.cc \
          anacompress.cc anadecompress.cc idl_read.cc imgtools.cc) \
        jsub/cfg.cc

OBJS := $(addprefix $(OBJ)/,$(subst /,__,$(SRCS:.cc=.o)))

all: $(OUT)/libsnapi_ref.so $(OUT)/ref_driver

$(OBJ)/%.o:
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c $(REF)/$(subst __,/,$*).cc -o $@

$(OUT)/libsnapi_ref.so: $(OBJS)
	$(CXX) -shared -o $@ $(OBJS) -lz -lpthread

$(OUT)/ref_driver: $(HERE)ref_driver.cc $(OUT)/libsnapi_ref.so
	$(CXX) $(CXXFLAGS) -o $@ $(HERE)ref_driver.cc -L$(OUT) -lsnapi_ref -Wl,-rpath,'$$ORIGIN' \
	  -Wl,--allow-shlib-undefined -lz -lpthread

clean:
	rm -rf $(OUT)

.PHONY: all clean
