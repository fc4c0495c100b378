// Tape recording and replay of API call streams, the reference's extension API (include/rgl/api/extensions/tape.h of
// RobotecGPULidar 0.19.0, same declarations and semantics).
#pragma once

#include <rgl/api/core.h>

// Starts recording every following core API call into `<path>.yaml` and `<path>.bin`
// (include/rgl/api/extensions/tape.h:26). Fails with RGL_TAPE_ERROR if a recording is active.
RGL_API rgl_status_t rgl_tape_record_begin(const char* path);

// Ends the active recording (tape.h:32). Fails with RGL_TAPE_ERROR if none is active.
RGL_API rgl_status_t rgl_tape_record_end();

// Reports whether a recording is active (tape.h:38).
RGL_API rgl_status_t rgl_tape_record_is_active(bool* is_active);

// Replays the tape at `path` (`<path>.yaml` + `<path>.bin`), each call at approximately its recorded time after the
// start (tape.h:45). Fails with RGL_TAPE_ERROR while recording, for an empty or outdated tape, or for an unknown call.
RGL_API rgl_status_t rgl_tape_play(const char* path);
