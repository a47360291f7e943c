// xtneighbor_cli.cpp -- drop-in for the reference's first, non-streaming
// command line `xtneighbor` (xtneighbor/src/cli.cu; deprecated upstream,
// readme.md:58-60), over libxtn.so.
//
// Flags, help text, messages, exit codes and the output file follow
// xtneighbor/src/cli.cu:6-172 and util.cu:3-27. That program keeps each
// sequence as its 18-residue packed form (Int3, xtneighbor/src/xtn.h:1-9;
// codec.cu:34-51 stops at 18 residues), finds every pair within the
// Levenshtein distance on those forms (no measure option) and writes the
// pairs in ascending order (the sort + unique + order-preserving filter of
// xtn_inner.cu:55-82). Here the sequences go to the engine cut to 18
// residues, which gives the same pairs. -c checks the output against a
// brute force over all pairs (brute_force.cu:43-79), here on host threads.
#include <clocale>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "xtn.h"

namespace {

enum ReturnCode { SUCCESS = 0, ERROR = 1, EXIT = 2 };  // xtneighbor/src/xtn.h:36

const int MAX_INPUT_LENGTH = 18;       // xtneighbor/src/xtn.h:1
const char VERSION[] = "1.0.0\n";      // cli.cu:6
const char HELP_TEXT[] =               // cli.cu:7-16
    "xt_neighbor\n"
    "\t description: perform xt_neighbor algorithm for near neighbor search of T cell receptor's CDR3 sequences\n"
    "\t -p or --input-path [str] (required): set the path of input file which is a text file containing one CDR3 sequence per line\n"
    "\t -n or --input-length [number] (required): set the number of sequences given in the input file\n"
    "\t -d or --distance [number]: set the distance threshold defining the neighbor\n"
    "\t -c or --check-output: check if the output is correct using brute force algorithm (very slow!)\n"
    "\t -o or --output-path [str]: set the path of the output file (default to no output)\n"
    "\t -v or --version: print the version of the program then exit\n"
    "\t -h or --help: print the help text of the program then exit\n"
    "\t -V or --verbose: print extra detail as the program runs for debugging purpose\n";

struct Args {  // XTNArgs, xtneighbor/src/xtn.h:19-28
	int distance = 1;
	int verbose = 0;
	char* seq1Path = nullptr;
	int seq1Len = 0;
	char* outputPath = nullptr;
	int checkOutput = 0;
};

int print_err(const char* s) {  // util.cu:3-8
	fprintf(stderr, "Error: %s\n", s);
	return ERROR;
}

// cli.cu:18-62
int parse_args(int argc, char** argv, Args* a) {
	for (int i = 1; i < argc; i++) {
		const char* cur = argv[i];
		auto next = [&]() -> char* { return i + 1 < argc ? argv[++i] : nullptr; };
		if (!strcmp(cur, "-v") || !strcmp(cur, "--version")) {
			printf("%s", VERSION);
			return EXIT;
		} else if (!strcmp(cur, "-h") || !strcmp(cur, "--help")) {
			printf("%s", HELP_TEXT);
			return EXIT;
		} else if (!strcmp(cur, "-V") || !strcmp(cur, "--verbose")) {
			a->verbose = 1;
		} else if (!strcmp(cur, "-c") || !strcmp(cur, "--check-output")) {
			a->checkOutput = 1;
		} else if (!strcmp(cur, "-p") || !strcmp(cur, "--input-path")) {
			a->seq1Path = next();
		} else if (!strcmp(cur, "-o") || !strcmp(cur, "--output-path")) {
			a->outputPath = next();
		} else if (!strcmp(cur, "-d") || !strcmp(cur, "--distance")) {
			const char* v = next();
			a->distance = v ? atoi(v) : 0;
			if (a->distance < 1)
				return print_err("distance must be a valid number larger than 0");
		} else if (!strcmp(cur, "-n") || !strcmp(cur, "--input-length")) {
			const char* v = next();
			a->seq1Len = v ? atoi(v) : 0;
			if (a->seq1Len == 0)
				return print_err("invalid input length");
		} else {
			return print_err("unknown option");
		}
	}
	if (a->seq1Path == nullptr)
		return print_err("missing path for seq1");
	if (a->seq1Len == 0)
		return print_err("missing length for seq1");
	return SUCCESS;
}

// codec.cu:9-21 (the residue code; -1 for anything else)
int char_encode(char c) {
	if (c < 'A' || c > 'Y')
		return -1;
	if (c == 'B' || c == 'J' || c == 'O' || c == 'U' || c == 'X')
		return -1;
	return c - '@';
}

// cli.cu:66-95: lines read 49 characters at a time, blank lines skipped, each
// line's first 18 residues kept (codec.cu:34-51: the packing stops at a line
// end or after 18 residues; an invalid residue before that rejects the line)
int parse_file(const char* path, std::vector<char>* chars, std::vector<uint32_t>* offs, int len) {
	FILE* file = fopen(path, "r");
	if (!file)
		return print_err("file reading failed");
	const int BUFFER_SIZE = 50;
	char line[BUFFER_SIZE];
	int lineNumber = 0, inputCount = 0;
	offs->assign(1, 0);
	while (fgets(line, BUFFER_SIZE, file)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		int n = 0;
		bool bad = false;
		for (; n < MAX_INPUT_LENGTH; n++) {
			const char c = line[n];
			if (c == '\0' || c == '\n')
				break;
			if (char_encode(c) == -1) {
				bad = true;
				break;
			}
		}
		// an empty packed form (the reference tests its first word) is an
		// error as well: a line holding only a NUL or starting a NUL run
		if (bad || n == 0) {
			fclose(file);
			char msg[160];
			snprintf(msg, sizeof(msg),
			         "parsing error at line: %d (only upper-cased amino acids with max length of %d are allowed)",
			         lineNumber, MAX_INPUT_LENGTH);
			return print_err(msg);
		}
		chars->insert(chars->end(), line, line + n);
		offs->push_back(offs->back() + (uint32_t)n);
		inputCount++;
	}
	fclose(file);
	if (inputCount != len)
		return print_err("input length doesn't match with the actual");
	return SUCCESS;
}

void print_args(const Args& a) {  // util.cu:18-27
	printf("XTNArgs{\n");
	printf("\tdistance: %d\n", a.distance);
	printf("\tverbose: %d\n", a.verbose);
	printf("\tseq1Len: %d\n", a.seq1Len);
	printf("\tseq1Path: \"%s\"\n", a.seq1Path ? a.seq1Path : "(null)");
	printf("\toutputPath: \"%s\"\n", a.outputPath ? a.outputPath : "(null)");
	printf("\tcheckOutput: %d\n", a.checkOutput);
	printf("}\n");
}

struct Triple {
	int x, y, d;
	bool operator==(const Triple& o) const { return x == o.x && y == o.y && d == o.d; }
};
struct TripleHash {
	size_t operator()(const Triple& t) const {
		return ((size_t)(uint32_t)t.x * 0x9E3779B97F4A7C15ull) ^ ((size_t)(uint32_t)t.y << 7) ^ (size_t)t.d;
	}
};

// Wagner-Fischer over the two strings (brute_force.cu:22-41)
int levenshtein(const char* a, int la, const char* b, int lb) {
	int col[MAX_INPUT_LENGTH + 1];
	for (int y = 0; y <= la; y++)
		col[y] = y;
	for (int x = 1; x <= lb; x++) {
		int diag = col[0];
		col[0] = x;
		for (int y = 1; y <= la; y++) {
			const int up = col[y];
			const int sub = diag + (a[y - 1] == b[x - 1] ? 0 : 1);
			int v = up + 1 < col[y - 1] + 1 ? up + 1 : col[y - 1] + 1;
			col[y] = v < sub ? v : sub;
			diag = up;
		}
	}
	return col[la];
}

// brute_force.cu:43-79: every pair i < j within the distance, compared with
// the output as sets (same size, every output triple among them, no repeats)
bool check_output(const std::vector<char>& chars, const std::vector<uint32_t>& offs, int distance,
                  const std::vector<Triple>& out) {
	const int n = (int)offs.size() - 1;
	int T = (int)std::thread::hardware_concurrency();
	T = T < 1 ? 1 : (T > 16 ? 16 : T);
	std::vector<std::vector<Triple>> part(T);
	std::vector<std::thread> th;
	for (int t = 0; t < T; t++)
		th.emplace_back([&, t] {
			for (int i = t; i < n; i += T)
				for (int j = i + 1; j < n; j++) {
					const int d = levenshtein(chars.data() + offs[i], (int)(offs[i + 1] - offs[i]),
					                          chars.data() + offs[j], (int)(offs[j + 1] - offs[j]));
					if (d <= distance)
						part[t].push_back({i, j, d});
				}
		});
	for (auto& x : th)
		x.join();
	std::unordered_set<Triple, TripleHash> answer;
	for (auto& p : part)
		answer.insert(p.begin(), p.end());
	if (out.size() != answer.size())
		return false;
	std::unordered_set<Triple, TripleHash> seen;
	for (const Triple& t : out) {
		if (!answer.count(t))
			return false;
		seen.insert(t);
	}
	return seen.size() == answer.size();
}

void collect(void* user, const xtn_output* o) {
	auto* v = static_cast<std::vector<Triple>*>(user);
	for (int64_t i = 0; i < o->len; i++)
		v->push_back({o->index_pairs[i].x, o->index_pairs[i].y, o->pairwise_distances[i]});
}

}  // namespace

// cli.cu:119-172
int main(int argc, char** argv) {
	Args args;
	int rc = parse_args(argc, argv, &args);
	if (rc != SUCCESS)
		return rc;
	std::vector<char> chars;
	std::vector<uint32_t> offs;
	rc = parse_file(args.seq1Path, &chars, &offs, args.seq1Len);
	if (rc != SUCCESS)
		return rc;
	if (args.verbose)
		print_args(args);
	setlocale(LC_ALL, "");

	xtn_config cfg;
	xtn_config_init(&cfg);
	cfg.distance = args.distance;
	cfg.measure = XTN_LEVENSHTEIN;
	xtn_handle* h = nullptr;
	int st = xtn_open(&h, &cfg);
	if (st != XTN_OK) {
		fprintf(stderr, "Error: cannot open the GPU engine (%s)\n", xtn_status_string(st));
		return ERROR;
	}
	std::vector<Triple> out;
	st = xtn_set_sequences(h, chars.data(), offs.data(), (int32_t)offs.size() - 1);
	if (st == XTN_OK)
		st = xtn_perform(h, nullptr, collect, &out);
	if (st != XTN_OK) {
		fprintf(stderr, "Error: %s\n", xtn_last_error(h));
		xtn_close(h);
		return ERROR;
	}
	xtn_close(h);

	if (args.outputPath) {  // cli.cu:100-114
		FILE* f = fopen(args.outputPath, "w");
		if (!f) {
			print_err("file reading failed");
			print_err("file writing failed");
		} else {
			for (const Triple& t : out)
				fprintf(f, "%d %d %d\n", t.x, t.y, t.d);
			fclose(f);
		}
	}
	if (args.checkOutput) {
		if (check_output(chars, offs, args.distance, out))
			printf("output is verrified to be correct\n");
		else
			print_err("input checking failed");
	}
	if (args.verbose)
		printf("Success! Number of triplet: %'zu\n", out.size());
	return SUCCESS;
}
