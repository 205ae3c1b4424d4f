/* File formats of the reference's front end, restated in C++:
 *   OVF 1.0 / 2.0 (OOMMF) vector fields  (mumax/omf/input.go:27-235, output.go:37-300)
 *   .tensor blobs                         (mumax/tensor/input.go:21-66, output.go:20-47)
 *   ODT tables                            (mumax/omf/odt.go:26-132)
 * Vector fields use the engine's internal layout: float[3][X][Y][Z], Z fastest,
 * internal X = user z (thickness), Y = user y, Z = user x; component 0 is the
 * internal X.  In OVF files nodes run with user x fastest and the three values
 * of a node are the user (x, y, z) components, i.e. internal (Z, Y, X).
 * Functions return 0 on success and a negative code on failure, with the
 * message available from mm_io_error(). */
#ifndef MUMAX_IO_H
#define MUMAX_IO_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mmOvfInfo {
    int size[3];          /* internal X, Y, Z (znodes, ynodes, xnodes) */
    double step[3];       /* internal cell size (zstepsize, ystepsize, xstepsize) */
    int version;          /* 1 or 2 */
    int binary;           /* 0: text, 4 or 8: binary bytes per value */
    double multiplier;    /* valuemultiplier (OVF 1.0; 1 if absent) */
    char valueunit[32];
    char meshunit[16];
    char title[64];
    double total_time;    /* OVF 2.0 desc "Total simulation time" (s), or 0 */
    double stage_time;    /* OVF 2.0 desc "Stage simulation time" (s), or 0 */
    int ndesc;            /* "Desc" key/value pairs in the header */
} mmOvfInfo;

const char* mm_io_error(void);

/* header of an OVF file (info->ndesc counts its Desc entries) */
int mm_ovf_info(const char* path, mmOvfInfo* info);
/* Desc entry i of the file: key and value, truncated to cap bytes each */
int mm_ovf_desc(const char* path, int i, char* key, char* value, int cap);
/* data of an OVF file into m[3][X][Y][Z] (sized from mm_ovf_info), as stored
 * (the valuemultiplier is not applied, as omf.Read) */
int mm_ovf_read(const char* path, float* m, mmOvfInfo* info);
/* write m[3][X][Y][Z]: info->version 1 (binary: 4-byte big-endian) or 2
 * (binary: 4-byte little-endian), info->binary 0 for text; ndesc Desc entries */
int mm_ovf_write(const char* path, const float* m, const mmOvfInfo* info, const char* const* desc_keys,
                 const char* const* desc_values, int ndesc);

/* .tensor: "#t1\n" magic, rank, sizes (int32 little-endian), float32 data.
 * With data == 0 only rank and size[0..rank) are returned (size holds 16). */
int mm_tensor_read(const char* path, int* rank, int* size, float* data);
int mm_tensor_write(const char* path, int rank, const int* size, const float* data);

/* ODT 1.0 table: header with title, units and columns, one row per call
 * (each row is flushed to the file: mumax/omf/odt.go:38-139) */
typedef struct mmOdt mmOdt;
mmOdt* mm_odt_open(const char* path, const char* title, int ncols, const char* const* names,
                   const char* const* units);
/* The same with ndesc "# Desc: key :  value" lines in a "# Begin: Header" block
 * before the table (Desc(), simmetadata.go:19-30, odt.go:60-68, 117-121). */
mmOdt* mm_odt_open_desc(const char* path, const char* title, int ncols, const char* const* names,
                        const char* const* units, int ndesc, const char* const* desc_keys,
                        const char* const* desc_values);
int mm_odt_row(mmOdt* t, const double* values);
int mm_odt_close(mmOdt* t);

#ifdef __cplusplus
}
#endif
#endif
