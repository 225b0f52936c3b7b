/*
 * ensc_deck.h - Abaqus .inp reader of the step engine (C ABI).
 *
 * Replaces the reference's DataIn::read_inp + transferToExDyna
 * (Explicit3D/source/dataIn.cpp:300-558, 1550-1788) and read_project_txt
 * (Explicit3D/source/exDyna3D.cpp:698-772): the same keyword dispatch, the
 * same set/instance/step semantics and the same quirks, producing an
 * ensc_model_desc (include/ensc_model.h) for ensc_engine_create.
 *
 * Extensions the reference ignores are off unless asked for (flags of the
 * _ex readers), so a deck read with flags 0 describes the problem the
 * reference solves:
 *   ENSC_DECK_PLASTIC       *Plastic: J2 with linear hardening (yield0 =
 *                           first yield stress, hard = slope between the
 *                           first two points); the reference has no such
 *                           keyword (dataIn.cpp:69-92);
 *   ENSC_DECK_VEL_AMPLITUDE AMPLITUDE= on *Boundary, type=VELOCITY; the
 *                           reference keeps velocity BCs constant
 *                           (BoundaryCondition.h:18, dataIn.cpp:1114-1195);
 *   ENSC_DECK_DSLOAD        *Dsload pressures on *Surface element faces; the
 *                           reference parses them (dataIn.cpp:1339-1478) and
 *                           never applies them.
 */
#ifndef ENSC_DECK_H
#define ENSC_DECK_H

#include <stddef.h>
#include <stdint.h>

#include "ensc_engine.h"
#include "ensc_model.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ensc_deck ensc_deck;

/* Solver parameters from project.txt (exDyna3D.cpp:698-772).  Defaults when
 * absent: c_dr 0.013, c_cr 0.05, factor 0.71. */
typedef struct ensc_project {
    double c_dr, c_cr, factor;
    char deck[1024];  /* line 2: deck file name */
} ensc_project;

ensc_status ensc_project_read(const char* path, ensc_project* out);

enum {
    ENSC_DECK_PLASTIC = 1u,
    ENSC_DECK_VEL_AMPLITUDE = 2u,
    ENSC_DECK_DSLOAD = 4u
};

/* Read a deck the way the reference's DataIn::read_inp does (no
 * extensions).  project may be NULL (defaults). */
ensc_status ensc_deck_read(const char* path, const ensc_project* project, ensc_deck** out);
/* Read a deck held in memory. */
ensc_status ensc_deck_parse(const char* text, size_t len, const ensc_project* project, ensc_deck** out);
/* The same with extensions enabled by flags (ENSC_DECK_*). */
ensc_status ensc_deck_read_ex(const char* path, const ensc_project* project, uint32_t flags, ensc_deck** out);
ensc_status ensc_deck_parse_ex(const char* text, size_t len, const ensc_project* project, uint32_t flags,
                               ensc_deck** out);
/* The parsed job.  Pointers stay valid until ensc_deck_free. */
ensc_status ensc_deck_model(const ensc_deck* deck, ensc_model_desc* out);
/* JSON description of what was parsed (material, parameters, node sets,
 * initial velocities, amplitudes, steps with their BC lines and gravity).
 * Returns the length needed (excluding NUL); writes up to cap bytes. */
int64_t ensc_deck_json(const ensc_deck* deck, char* buf, int64_t cap);
ensc_status ensc_deck_free(ensc_deck* deck);

#ifdef __cplusplus
}
#endif
#endif /* ENSC_DECK_H */
