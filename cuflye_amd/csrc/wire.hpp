// wire.hpp - canonical text of candidate-record-v1 streamed through SHA-256
// (host code).
//
// candidate-record-v1 (docs/abi/candidate-record-v1.md:31-68) compares dumps
// by the SHA-256 of their canonical text: every record as
// "query_id\tquery_pos\tkmer\ttarget_id\ttarget_pos\t+|-\n", records in
// ascending numeric tuple order (tools/validate_candidate_dump.py:76-134).
// A query record's matches arrive in Flye's emission order - window position
// ascending, bucket order inside a window - and all matches of one window
// share query_pos and kmer, so the canonical order is: query ids ascending,
// then per window its matches sorted by (target_id, target_pos).  The hasher
// takes query records in ascending id order, sorts inside windows, formats in
// parallel and feeds SHA-256 in order: no sort of the whole dump and no dump
// held in memory.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <memory>
#include <string>

#include "../../include/flyehip.h"

namespace fh
{

class CanonSha256
{
public:
	CanonSha256();
	~CanonSha256();
	// Appends query records whose matches are recs[off[i] .. off[i + 1]) in
	// emission order, i in [0, nQueries); query ids must ascend across calls.
	// recs is reordered inside windows.
	void addQueries(flyehip_candidate_t* recs, const int64_t* off, int32_t nQueries, int threads);
	// Finishes the digest (32 bytes); returns false on a library error.
	bool finish(uint8_t* out);
	int64_t records() const { return records_; }
	const std::string& error() const { return err_; }

private:
	struct Impl;
	std::unique_ptr<Impl> impl_;
	int64_t records_ = 0;
	std::string err_;
};

} // namespace fh
