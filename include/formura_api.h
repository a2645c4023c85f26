/*
 * formura_api.h -- the C API every generated Formura library exports.
 *
 * The concrete declarations are generated per program (NAME.h, next to
 * NAME.hip), because the state struct and the Navi struct depend on the
 * program's fields, axes and grid.  They have exactly the layout and names
 * the reference's generator emits, so a main.c written for the reference
 * compiles and links unchanged.  This file documents each entry point and the
 * reference interface it replaces; it is not meant to be included.
 *
 *   typedef struct { T f[N1]..[Nd]; ... } Formura_Grid_Struct;
 *       fields in Data.Map key order            -- Templates.hs:53-73
 *   typedef struct { int time_step; int lower_a..; int upper_a..;
 *                    double space_interval_a..; int my_rank; int offset_a..;
 *                    double length_a..; int total_grid_a..; } Formura_Navi;
 *                                                -- Templates.hs:75-115 (no MPI)
 *   extern Formura_Grid_Struct formura_data;    -- Templates.hs:53-61, Encode.hs:44-45
 *   double to_pos_<a>(int i, Formura_Navi n);   -- Templates.hs:221-225
 *       n.space_interval_<a>*((i+n.offset_<a>)%n.total_grid_<a>)
 *   void Formura_Init(int *argc, char ***argv, Formura_Navi *n);
 *                                                -- Templates.hs:227-250
 *       sets every Navi field, runs the init graph on the host into
 *       formura_data, allocates the device state and uploads it.
 *   void Formura_Forward(Formura_Navi *n);      -- Templates.hs:295-334
 *       temporal_blocking_interval steps (1 without blocking) on the GPU;
 *       offset_<a> -= sleeve*steps (mod total), time_step += steps, filter
 *       hook as in the reference; formura_data is uploaded before and
 *       downloaded after unless the library is in resident mode.
 *   void Formura_Finalize();                    -- Templates.hs:383-386
 *       downloads a resident state into formura_data, frees device memory.
 *
 * Not exported (the library drives one GPU): Formura_Custom_Init(Formura_Navi *,
 * MPI_Comm) (Templates.hs:230-231) and the MPI Navi fields; a config with
 * mpi_shape is rejected by the generator.
 *
 * Errors follow the reference: message on stderr, exit(1) (Templates.hs:243).
 * The device-resident extension is declared in formura_hip.h.
 */
#error "formura_api.h documents the generated API; include the generated NAME.h instead"
