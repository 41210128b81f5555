// Lattice extents, as in the reference's configure_file output (include/config.h.in):
// a caller defines NS/NT (e.g. -DNS=512 -DNT=512) or edits this file.
#ifndef CONFIG_H
#define CONFIG_H
#ifndef NS
#define NS 64
#endif
#ifndef NT
#define NT 64
#endif
#endif
