/* HDF5 data formats around the HiDEM timestep: the reference's checkpoint
 * file and its particle input file.  Host-only C ABI of libhidem_io.so (a
 * library of its own so the compute library does not depend on HDF5); the
 * host structs are those of hdm_abi.h.
 *
 * Checkpoint layout (HiDEMH5Checkpoint::writeCheckpoint, HiDEMH5Checkpoint.h
 * :23-49; H5Checkpoint.h; ParticleDataHandler::checkpointData,
 * ParticleDataHandler.h:215-231; LinkDataHandler::checkpointData,
 * LinkDataHandler.h:356-393), one rank:
 *   /                attribute timeStep (int)
 *   /Domains         attributes domainLower/Upper[Halo]X/Y/Z (double[1]),
 *                    domainOwner (int[1]): the particles' bounding box
 *   /Particles       field_0..15 = xCord yCord zCord xTheta yTheta zTheta
 *                    xDiff yDiff zDiff uDiff vDiff wDiff (double) radius
 *                    (double) owners (int) type (uint8) globalKeys (int64);
 *                    attributes spann_sizes, spann_offsets (int64[1])
 *   /Links           field_0..6 = owners (int) state (uint8) globalKeys,
 *                    globalKeyEndA, globalKeyEndB (int64) initialLen
 *                    (double) particleLinkNum (int); spans; hdm_maxLoad
 *                    (double, the per-link threshold, when present)
 *   /LinkSideData    side_field_7..12 = CT0..CT5 (double, 2 per link: end A
 *                    then end B; the names continue the link field count as
 *                    the reference writes them); spans
 * Particle global keys are the indices; a link's key is A * 256 + its rank
 * among the links of A (LinkGenerator.h:19,97,171).
 *
 * Every function returns HDM_OK (0) or a negative HDM_E_* code;
 * hdm_io_last_error() describes the last failure of the calling thread. */
#ifndef HIDEM_HDM_IO_H
#define HIDEM_HDM_IO_H

#include <stdint.h>

#include "hdm_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *hdm_io_last_error(void);

/* Writes a checkpoint of one rank's particles and links (links may be NULL).
 * Replaces the file. */
int hdm_checkpoint_write(const char *path, int32_t timestep, const hdm_particles_host *p, const hdm_links_host *l);

/* Sizes of a checkpoint (summed over the ranks that wrote it) and its step. */
int hdm_checkpoint_info(const char *path, int64_t *nParticles, int64_t *nLinks, int32_t *timestep);

/* Reads a checkpoint into caller arrays sized by hdm_checkpoint_info (p->n,
 * l->n must match; the force fields and l->maxLoad may be NULL; forces are
 * not in the file and read as 0).  The ranks' spans are concatenated in rank
 * order; link ends are found from the particles' global keys, so a link to a
 * key that is not in the file is an error (as the reference's
 * connectParticleLinks finds it dangling). */
int hdm_checkpoint_read(const char *path, hdm_particles_host *p, hdm_links_host *l);

/* Particle input file of [lattice] type = "File" (FileParticleGenerator,
 * FileParticleGenerator.h:16-101; HiDEMGeometry.cpp:111-121): datasets named
 * xName, yName, zName, rName (NULL: "xCord", "yCord", "zCord", "radius") at
 * the file's root, read as float as the reference reads them.  Count first,
 * then the values into p (coordinates and radius; rotations, Diff and forces
 * zeroed where given, type 0). */
int hdm_particle_file_count(const char *path, const char *xName, int64_t *n);
int hdm_particle_file_read(const char *path, const char *xName, const char *yName, const char *zName,
                           const char *rName, hdm_particles_host *p);

#ifdef __cplusplus
}
#endif

#endif
