# Builds the reference's own parameter reader (read_param.c, confirm_param.c,
# msg.c, compiled where they lie in the reference tree) with the harness
# ref_param_main.c into oracle/_ref/ref_params.  MPICH's headers and library
# are taken from /opt/conda (its mpicc wrapper has no compiler behind it).
# The rest of the reference needs FFTW3, GSL and generated sources and is
# not built.
REF      ?= /root/reference
MPI_DIR  ?= /opt/conda
OUT      := _ref
CC       := gcc

all: $(OUT)/ref_params

$(OUT):
	mkdir -p $(OUT)

$(OUT)/ref_params: ref_param_main.c $(REF)/read_param.c $(REF)/confirm_param.c $(REF)/msg.c | $(OUT)
	$(CC) -std=gnu99 -O1 -w -I$(REF) -I$(MPI_DIR)/include -o $@ ref_param_main.c \
	    $(REF)/read_param.c $(REF)/confirm_param.c $(REF)/msg.c \
	    -L$(MPI_DIR)/lib -Wl,-rpath,$(MPI_DIR)/lib -lmpi

clean:
	rm -f $(OUT)/ref_params

.PHONY: all clean
