// xt_neighbor_cli.cpp -- drop-in xt_neighbor command line over libxtn.so.
//
// Flags, help text, input parsing (CSV, info CSV, AIRR TSV), error messages,
// exit codes (0 success, 1 error, 2 help/version) and output formats follow
// xtneighbor_streaming/src/cli.cu:10-277 and airr.cu:9-151. Input is parsed
// and validated on the host before the GPU is touched, so every parse error
// path runs without a device.
#include <climits>
#include <clocale>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "xtn.h"

namespace {

enum ReturnCode { SUCCESS = 0, ERROR = 1, EXIT = 2 };  // xtn.h:93

const char VERSION[] = "2.0.0\n";  // cli.cu:9
const char HELP_TEXT[] =           // cli.cu:10-27
    "xt_neighbor: perform either nearest neighbor search for CDR3 sequences or immune repertoire overlap using GPU-based xt_neighbor algorithm.\n"
    "\t====================\n\t Common Options\n\t====================\n"
    "\t -d or --distance [number]: distance threshold defining the neighbor (default to 1)\n"
    "\t -o or --output-path [str]: path of the output file (default to no output)\n"
    "\t -m or --measurement [leven|hamming]: distance measurement (default to leven)\n"
    "\t -v or --version: print the version of the program then exit\n"
    "\t -h or --help: print the help text of the program then exit\n"
    "\t -V or --verbose: print extra detail as the program runs for debugging purpose\n"
    "\t -a or --airr: use AIRR format for input-path instead. Relevant fields are cdr3_aa and duplicate_count\n"
    "\t====================\n\t Nearest Neighbor Options\n\t====================\n"
    "\t -i or --input-path [str] (required): path of csv input file containing exactly 1 column: CDR3 amino acid sequences\n"
    "\t -n or --input-length [number] (required): number of rows given in the input file\n"
    "\t====================\n\t Repertoire Overlap Options\n\t====================\n"
    "\t -i or --input-path [str] (required): path of csv input file containing exactly 2 columns: CDR3 amino acid sequences and their frequency. Note that the sequences are assumed to be unique\n"
    "\t -n or --input-length [number] (required): number of sequences given in the input file\n"
    "\t -I or --info-path [str] (required): path of csv input file containing exactly 1 column: repertoire sizes. Note that the order of input sequence must be sorted according to this repertoire info\n"
    "\t -N or --info-length [number] (required): number of repertoires given in the info file\n";

struct Args {  // XTNArgs, xtn.h:50-60
	int distance = 1;
	int verbose = 0;
	char* seqPath = nullptr;
	int seqLen = 0;
	char* outputPath = nullptr;
	int measure = XTN_LEVENSHTEIN;
	char* infoPath = nullptr;
	int infoLen = 0;
	int airr = 0;
};

int print_err(const char* s) {  // util.cu:13-16
	fprintf(stderr, "Error: %s\n", s);
	return ERROR;
}

int print_err_line(const char* s, int line) {  // util.cu:18-22
	fprintf(stderr, "Error: %s\n", s);
	fprintf(stderr, "Error line: %d\n", line);
	return ERROR;
}

// codec.cu:25-38
int char_encode(char c) {
	if (c < 'A' || c > 'Y')
		return -1;
	if (c == 'B' || c == 'J' || c == 'O' || c == 'U' || c == 'X')
		return -1;
	return c - '@';
}

// SeqArray (xtn.h:95-114, codec.cu:168-196): concatenated residues + offsets
struct SeqStore {
	std::vector<char> chars;
	std::vector<uint32_t> offsets{0};
	int maxSize = 0;
	int size() const { return (int)offsets.size() - 1; }

	// codec.cu:175-196: a field ends at NUL, LF, ',' or TAB, at most 500
	// residues; an empty field is accepted and stores nothing
	int append(const char* s) {
		if (size() >= maxSize)
			return print_err("SeqArray exceeding max length");
		int i = 0;
		while (i < 500) {
			char c = s[i];
			if (c == '\0' || c == '\n' || c == ',' || c == '\t')
				break;
			i++;
			if (char_encode(c) == -1)
				return print_err("invalid char");
			chars.push_back(c);
		}
		if (i == 0)
			return SUCCESS;
		offsets.push_back(offsets.back() + (uint32_t)i);
		return SUCCESS;
	}
};

// cli.cu:29-89
int parse_args(int argc, char** argv, Args* a) {
	for (int i = 1; i < argc; i++) {
		const char* cur = argv[i];
		auto value = [&](void) -> char* { return i + 1 < argc ? argv[++i] : nullptr; };
		if (!strcmp(cur, "-v") || !strcmp(cur, "--version")) {
			printf("%s", VERSION);
			return EXIT;
		} else if (!strcmp(cur, "-h") || !strcmp(cur, "--help")) {
			printf("%s", HELP_TEXT);
			return EXIT;
		} else if (!strcmp(cur, "-V") || !strcmp(cur, "--verbose")) {
			a->verbose = 1;
		} else if (!strcmp(cur, "-a") || !strcmp(cur, "--airr")) {
			a->airr = 1;
		} else if (!strcmp(cur, "-i") || !strcmp(cur, "--input-path")) {
			a->seqPath = value();
		} else if (!strcmp(cur, "-I") || !strcmp(cur, "--info-path")) {
			a->infoPath = value();
		} else if (!strcmp(cur, "-o") || !strcmp(cur, "--output-path")) {
			a->outputPath = value();
		} else if (!strcmp(cur, "-d") || !strcmp(cur, "--distance")) {
			char* v = value();
			if (!v)
				return print_err("distance must be a valid number larger than 0");
			a->distance = atoi(v);
			if (a->distance < 1)
				return print_err("distance must be a valid number larger than 0");
		} else if (!strcmp(cur, "-n") || !strcmp(cur, "--input-length")) {
			char* v = value();
			a->seqLen = v ? atoi(v) : 0;
			if (a->seqLen == 0)
				return print_err("invalid input length");
		} else if (!strcmp(cur, "-N") || !strcmp(cur, "--info-length")) {
			char* v = value();
			a->infoLen = v ? atoi(v) : 0;
			if (a->infoLen == 0)
				return print_err("invalid info length");
		} else if (!strcmp(cur, "-m") || !strcmp(cur, "--measurement")) {
			char* m = value();
			if (m && !strcmp(m, "leven"))
				a->measure = XTN_LEVENSHTEIN;
			else if (m && !strcmp(m, "hamming"))
				a->measure = XTN_HAMMING;
			else
				return print_err("invalid measure option");
		} else {
			return print_err("unknown option");
		}
	}
	if (a->seqPath == nullptr)
		return print_err("missing path for seq");
	if (a->seqLen == 0)
		return print_err("missing length for seq");
	if ((a->infoPath == nullptr) != (a->infoLen == 0))
		return print_err("repertiore path or repertoire count is missing in overlap mode");
	return SUCCESS;
}

// cli.cu:101-147
int parse_input(const char* path, SeqStore* seqs, std::vector<xtn_seq_info>* info, int len,
                bool doubleCol) {
	FILE* f = fopen(path, "r");
	if (!f)
		return print_err("input file reading failed");
	const int BUFFER_SIZE = 500;
	char line[BUFFER_SIZE];
	int lineNumber = 1, inputCount = 0;
	if (!fgets(line, BUFFER_SIZE, f))
		line[0] = '\0';  // header
	while (fgets(line, BUFFER_SIZE, f)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		if (seqs->append(line) == ERROR) {
			fclose(f);
			return print_err_line("input parsing error", lineNumber);
		}
		if (doubleCol) {
			char* line2 = strchr(line, ',');
			if (!line2) {
				fclose(f);
				return print_err_line("input parsing error (comma expected)", lineNumber);
			}
			long temp = strtol(line2 + 1, nullptr, 10);
			if (temp == 0 || temp > INT_MAX || temp < INT_MIN) {
				fclose(f);
				return print_err_line("input parsing error (invalid number)", lineNumber);
			}
			if (inputCount < (int)info->size())
				(*info)[inputCount].frequency = (int32_t)temp;
		}
		inputCount++;
	}
	fclose(f);
	// a line whose first field is empty is counted without storing a
	// sequence; the reference then runs on uninitialised rows, here it is
	// reported as a length mismatch
	if (inputCount != len || seqs->size() != len)
		return print_err("input length doesn't match with the actual");
	return SUCCESS;
}

// cli.cu:152-193
int parse_info(const char* path, std::vector<xtn_seq_info>* result, int len, int seqLen) {
	FILE* f = fopen(path, "r");
	if (!f)
		return print_err("info file reading failed");
	const int BUFFER_SIZE = 500;
	char line[BUFFER_SIZE];
	int lineNumber = 1, inputCount = 0, resultIndex = 0;
	if (!fgets(line, BUFFER_SIZE, f))
		line[0] = '\0';
	while (fgets(line, BUFFER_SIZE, f)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		long temp = strtol(line, nullptr, 10);
		if (temp == 0 || temp > INT_MAX || temp < INT_MIN) {
			fclose(f);
			return print_err_line("info parsing error (invalid number)", lineNumber);
		}
		if (resultIndex + temp > seqLen) {
			fclose(f);
			return print_err("total repertoires' size does not match sequence count [1]");
		}
		for (long i = 0; i < temp; i++) {
			(*result)[resultIndex].repertoire = inputCount;
			(*result)[resultIndex].original_index = resultIndex;
			resultIndex++;
		}
		inputCount++;
	}
	fclose(f);
	if (inputCount != len)
		return print_err("info length doesn't match with the actual");
	if (resultIndex != seqLen)
		return print_err("total repertoires' size does not match sequence count [2]");
	return SUCCESS;
}

// airr.cu:24-36: column start offsets of a TAB separated line
int split_offset(const char* str, int* out, int nColumn) {
	char c;
	int i = 0, count = 1;
	out[0] = 0;
	while ((c = str[i++]) != '\0') {
		if (c == '\t') {
			out[count++] = i;
			if (count == nColumn)
				return count;
		}
	}
	return count;
}

// airr.cu:41-48: the column whose start is the first occurrence of needle
int find_idx(const char* haystack, const char* needle, const int* offsets, int n) {
	const char* p = strstr(haystack, needle);
	if (!p)
		return -1;
	int pos = (int)(p - haystack);
	for (int i = 0; i < n; i++)
		if (offsets[i] == pos)
			return i;
	return -1;
}

const int DEFAULT_MAX_COLUMN = 300;  // airr.cu:11
const int DEFAULT_LINE_SIZE = 5000;  // airr.cu:13

// airr.cu:53-66
int find_header(const char* line, int& cdrIndex, int& freqIndex, bool doubleCol) {
	std::vector<int> offsets(DEFAULT_MAX_COLUMN);
	int n = split_offset(line, offsets.data(), DEFAULT_MAX_COLUMN);
	cdrIndex = find_idx(line, "cdr3_aa", offsets.data(), n);
	freqIndex = doubleCol ? find_idx(line, "duplicate_count", offsets.data(), n) : -1;
	return n;
}

// airr.cu:71-105. A header without the needed column (index -1) makes every
// row an error here; the reference indexes offsets[-1] in that case.
int extract_data(const char* line, int cdrIndex, int freqIndex, int nColumn, int lineNumber,
                 SeqStore* seqs, int& freqOut, bool doubleCol) {
	std::vector<int> offsets(nColumn > 0 ? nColumn : 1);
	int n = split_offset(line, offsets.data(), nColumn);
	if (cdrIndex < 0 || n <= cdrIndex)
		return ERROR;
	if (seqs->append(line + offsets[cdrIndex]) == ERROR)
		return print_err_line("invalid seq", lineNumber);
	if (doubleCol) {
		if (freqIndex < 0 || n <= freqIndex)
			return ERROR;
		long temp = strtol(line + offsets[freqIndex], nullptr, 10);
		if (temp == 0 || temp > INT_MAX || temp < INT_MIN)
			return print_err_line("invalid number", lineNumber);
		freqOut = (int)temp;
	} else {
		freqOut = -1;
	}
	return SUCCESS;
}

// airr.cu:116-151
int parse_airr_input(const char* path, SeqStore* seqs, std::vector<xtn_seq_info>* info, int len,
                     bool doubleCol) {
	FILE* f = fopen(path, "r");
	if (!f)
		return print_err("input file reading failed");
	std::vector<char> buf(DEFAULT_LINE_SIZE);
	char* line = buf.data();
	int lineNumber = -1, inputCount = 0, seqIndex = -1, freqIndex = -1;
	if (!fgets(line, DEFAULT_LINE_SIZE, f))
		line[0] = '\0';
	int nColumn = find_header(line, seqIndex, freqIndex, doubleCol);
	while (fgets(line, DEFAULT_LINE_SIZE, f)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		int freq = 0;
		if (extract_data(line, seqIndex, freqIndex, nColumn, lineNumber, seqs, freq, doubleCol) != SUCCESS) {
			fclose(f);
			return print_err_line("line parsing error", lineNumber);
		}
		if (doubleCol && inputCount < (int)info->size())
			(*info)[inputCount].frequency = freq;
		inputCount++;
	}
	fclose(f);
	if (inputCount != len || seqs->size() != len)
		return print_err("input length doesn't match with the actual");
	return SUCCESS;
}

// util.cu:40-50
void print_args(const Args& a) {
	printf("XTNArgs{\n");
	printf("\tdistance: %d\n", a.distance);
	printf("\tverbose: %d\n", a.verbose);
	printf("\tseqLen: %'d\n", a.seqLen);
	printf("\tseqPath: \"%s\"\n", a.seqPath ? a.seqPath : "(null)");
	printf("\toutputPath: \"%s\"\n", a.outputPath ? a.outputPath : "(null)");
	printf("\tmeasure: \"%s\"\n", (a.measure == XTN_LEVENSHTEIN) ? "leven" : "hamming");
	printf("\tinfoPath: \"%s\"\n", a.infoPath ? a.infoPath : "(null)");
	printf("\tinfoLen: %'d\n", a.infoLen);
	printf("}\n");
}

struct Sink {
	FILE* out = nullptr;
	size_t total = 0;
	std::vector<char> buf;
};

// writes "%d %d %d\n" per pair (cli.cu:199-206) with an integer formatter
// and one fwrite per chunk
char* put_int(char* p, int64_t v) {
	char tmp[24];
	int n = 0;
	uint64_t u = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
	do {
		tmp[n++] = (char)('0' + u % 10);
		u /= 10;
	} while (u);
	if (v < 0)
		*p++ = '-';
	while (n)
		*p++ = tmp[--n];
	return p;
}

void nn_handler(void* user, const xtn_output* o) {
	Sink* s = static_cast<Sink*>(user);
	s->total += (size_t)o->len;
	if (!s->out)
		return;
	const size_t need = (size_t)o->len * 40 + 64;
	if (s->buf.size() < need)
		s->buf.resize(need);
	char* p = s->buf.data();
	for (int64_t i = 0; i < o->len; i++) {
		p = put_int(p, o->index_pairs[i].x);
		*p++ = ' ';
		p = put_int(p, o->index_pairs[i].y);
		*p++ = ' ';
		p = put_int(p, o->pairwise_distances[i]);
		*p++ = '\n';
	}
	fwrite(s->buf.data(), 1, (size_t)(p - s->buf.data()), s->out);
}

// cli.cu:208-215: "%d %d %'lu\n" under the process locale
void overlap_handler(void* user, const xtn_output* o) {
	Sink* s = static_cast<Sink*>(user);
	s->total += (size_t)o->len;
	if (!s->out)
		return;
	for (int64_t i = 0; i < o->len; i++)
		fprintf(s->out, "%d %d %'lu\n", o->index_pairs[i].x, o->index_pairs[i].y,
		        (unsigned long)o->pairwise_frequencies[i]);
}

}  // namespace

int main(int argc, char** argv) {
	Args args;
	setlocale(LC_ALL, "");  // cli.cu:234
	int rc = parse_args(argc, argv, &args);
	if (rc != SUCCESS)
		return rc;

	const bool overlapMode = args.infoPath != nullptr;
	SeqStore seqs;
	seqs.maxSize = args.seqLen > 0 ? args.seqLen : 0;
	if (args.seqLen > 0) {
		seqs.chars.reserve((size_t)args.seqLen * 18);
		seqs.offsets.reserve((size_t)args.seqLen + 1);
	}
	std::vector<xtn_seq_info> info;
	if (overlapMode) {
		info.assign(args.seqLen > 0 ? (size_t)args.seqLen : 0, xtn_seq_info{0, 0, 0});
		rc = parse_info(args.infoPath, &info, args.infoLen, args.seqLen);
		if (rc != SUCCESS)
			return rc;
	}
	if (args.airr)
		rc = parse_airr_input(args.seqPath, &seqs, &info, args.seqLen, overlapMode);
	else
		rc = parse_input(args.seqPath, &seqs, &info, args.seqLen, overlapMode);
	if (rc != SUCCESS)
		return rc;
	if (args.verbose)
		print_args(args);

	xtn_config cfg;
	xtn_config_init(&cfg);
	cfg.distance = args.distance;
	cfg.measure = args.measure;
	cfg.verbose = args.verbose;
	if (const char* e = getenv("XTN_MAX_RECORDS_PER_CHUNK"))
		cfg.max_records_per_chunk = atoll(e);
	xtn_handle* h = nullptr;
	int st = xtn_open(&h, &cfg);
	if (st != XTN_OK) {
		fprintf(stderr, "Error: cannot open the GPU engine (%s)\n", xtn_status_string(st));
		return ERROR;
	}
	st = xtn_set_sequences(h, seqs.chars.data(), seqs.offsets.data(), seqs.size());
	if (st != XTN_OK) {
		fprintf(stderr, "Error: %s\n", xtn_last_error(h));
		xtn_close(h);
		return ERROR;
	}
	Sink sink;
	if (args.outputPath) {
		sink.out = fopen(args.outputPath, "w");
		if (!sink.out) {
			xtn_close(h);
			return print_err("output file opening failed");
		}
	}
	st = xtn_perform(h, overlapMode ? info.data() : nullptr, overlapMode ? overlap_handler : nn_handler,
	                 &sink);
	if (sink.out)
		fclose(sink.out);
	if (st != XTN_OK) {
		fprintf(stderr, "Error: %s\n", xtn_last_error(h));
		xtn_close(h);
		return ERROR;
	}
	xtn_close(h);
	if (args.verbose)
		printf("total output length: %'lu\n", (unsigned long)sink.total);
	return SUCCESS;
}
